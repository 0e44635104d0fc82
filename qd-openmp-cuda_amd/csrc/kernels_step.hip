// kernels_step.hip - pointwise and reduction kernels of the split-operator step.
//
//   k_prephase  the V(-dt/2) start phase of cudaOpenMP2.cu:344 (steps == 0),
//               applied once before the first step
//   k_reduce  fixed-order reduction of the per-workgroup partial sums into one
//             observable record per step; flags the device step counters for
//             advance (replaces the host-synchronous cublasZdotc calls and
//              print_energies, cudaOpenMP2.cu:425-445)
//
// All reductions are in a fixed order, so an FP64 run is bitwise
// reproducible and a job's results do not depend on which other jobs share
// the launch.
#include <algorithm>

#include "qd_internal.h"
#include "qd_complex.h"
#include "step_reduce.h"

namespace qd {

// ---------------------------------------------------------------------------
// block reductions (wave sums + LDS across the 4 waves)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) { return wave_sum64(v); }

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
// Step-0 pre-phase: psi_b <- exp(+i dt/2 V) psi_b for every block of one job.
// The reference applies V(-dt/2) inside the first step when steps == 0
// (cudaOpenMP2.cu:344); the phase is pointwise like the absorber and V(dt)
// that the forward Legendre prologue applies, so running it once just before
// step 0 gives the same state and keeps the step kernels free of the branch.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_prephase(double2 *psi, const double *pot, long G, int nO, double half_dt) {
  for (long idx = (long)blockIdx.x * kBlock + threadIdx.x; idx < G; idx += (long)gridDim.x * kBlock) {
    double s, c;
    sincos(half_dt * pot[idx], &s, &c);
    const double2 t = make_double2(c, s);
    for (int b = 0; b < nO; b++) psi[b * G + idx] = cmul(t, psi[b * G + idx]);
  }
}

void launch_prephase(double2 *psi, const double *pot, long G, int nO, double dt, hipStream_t stream) {
  const long blocks = std::min<long>((G + kBlock - 1) / kBlock, 4096);
  hipLaunchKernelGGL(k_prephase, dim3((unsigned)blocks), dim3(kBlock), 0, stream, psi, pot, G, nO, 0.5 * dt);
}

// ---------------------------------------------------------------------------
// Potential phase with the absorber folded in (evolutionCUDAaux.cu:5-11,
// :243-252): vph[idx] = d1[i] d2[j] exp(-i dt V[idx]) at idx = i + n1 (j + n2 k),
// built once per job from the uploaded potential (d1 = d2 = nullptr: no
// absorber).  On the device, a 256 x 192 x 96 table takes microseconds
// instead of a host pass of 4.7M sine/cosine pairs per admitted job.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_vphase(double2 *vph, const double *pot, const double *d1,
                                                   const double *d2, int n1, int n2, long G, double dt) {
  for (long idx = (long)blockIdx.x * kBlock + threadIdx.x; idx < G; idx += (long)gridDim.x * kBlock) {
    double s, c;
    sincos(-dt * pot[idx], &s, &c);
    double dd = 1.0;
    if (d1) dd = d1[idx % n1] * d2[(idx / n1) % n2];
    vph[idx] = make_double2(dd * c, dd * s);
  }
}

void launch_vphase(double2 *vph, const double *pot, const double *d1, const double *d2, int n1, int n2, long G,
                   double dt, hipStream_t stream) {
  const long blocks = std::min<long>((G + kBlock - 1) / kBlock, 4096);
  hipLaunchKernelGGL(k_vphase, dim3((unsigned)blocks), dim3(kBlock), 0, stream, vph, pot, d1, d2, n1, n2, G, dt);
}

// Quadrature-weighted potential of the potential-energy epilogue:
// wpot[k M + m] = w_k V(k, m) for k < nt, zero rows up to `rows`.
__global__ __launch_bounds__(kBlock) void k_wpot(double *wpot, const double *pot, const double *w, int nt, long M,
                                                 long total) {
  for (long idx = (long)blockIdx.x * kBlock + threadIdx.x; idx < total; idx += (long)gridDim.x * kBlock) {
    const long k = idx / M;
    wpot[idx] = k < nt ? w[k] * pot[idx] : 0.0;
  }
}

void launch_wpot(double *wpot, const double *pot, const double *w, int nt, long M, int rows, hipStream_t stream) {
  const long total = (long)rows * M;
  const long blocks = std::min<long>((total + kBlock - 1) / kBlock, 4096);
  hipLaunchKernelGGL(k_wpot, dim3((unsigned)blocks), dim3(kBlock), 0, stream, wpot, pot, w, nt, M, total);
}

// ---------------------------------------------------------------------------
// Final reduction: one workgroup per (job, block, field), see reduce_unit.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_reduce(StepArgs a) {
  __shared__ double red[4];
  reduce_unit(a, blockIdx.x, red);
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

// HBM ceiling probes.  Each thread moves 8 values of 16 bytes per trip, all
// loads issued before the first store; every workgroup walks contiguous
// chunks of 8 x 256 values.  k_copy_wide writes with nontemporal stores;
// k_read_wide only reads (one partial sum per workgroup is written, so the
// loads cannot be dropped).
constexpr int kWide = 8;

__global__ __launch_bounds__(kBlock) void k_copy_wide(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                      long n) {
  const long chunk = (long)kWide * kBlock;
  for (long c0 = (long)blockIdx.x * chunk; c0 < n; c0 += (long)gridDim.x * chunk) {
    double2 v[kWide];
#pragma unroll
    for (int k = 0; k < kWide; k++) {
      const long q = c0 + (long)k * kBlock + threadIdx.x;
      v[k] = q < n ? src[q] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int k = 0; k < kWide; k++) {
      const long q = c0 + (long)k * kBlock + threadIdx.x;
      typedef double v2d_t __attribute__((ext_vector_type(2)));
      if (q < n) __builtin_nontemporal_store(v2d_t{v[k].x, v[k].y}, (v2d_t *)&dst[q]);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_read_wide(const double2 *__restrict__ src, double *__restrict__ out,
                                                      long n) {
  __shared__ double red[4];
  const long chunk = (long)kWide * kBlock;
  double acc = 0.0;
  for (long c0 = (long)blockIdx.x * chunk; c0 < n; c0 += (long)gridDim.x * chunk) {
    double2 v[kWide];
#pragma unroll
    for (int k = 0; k < kWide; k++) {
      const long q = c0 + (long)k * kBlock + threadIdx.x;
      v[k] = q < n ? src[q] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int k = 0; k < kWide; k++) acc += v[k].x + v[k].y;
  }
  acc = wave_sum64(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// kind 1: k_copy_wide (src -> dst), 2: k_read_wide (src, partial sums into dst);
// 2048 workgroups: 8 per CU
void launch_bandwidth(int kind, const double2 *src, double2 *dst, long n, hipStream_t stream) {
  if (kind == 2) hipLaunchKernelGGL(k_read_wide, dim3(2048), dim3(kBlock), 0, stream, src, (double *)dst, n);
  else hipLaunchKernelGGL(k_copy_wide, dim3(2048), dim3(kBlock), 0, stream, src, dst, n);
}

}  // namespace qd
