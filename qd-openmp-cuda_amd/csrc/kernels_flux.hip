// kernels_flux.hip - reaction probabilities from the flux through the
// dividing surface r2 = r_s.
//
// The reference parses the CRP struct (matlabStructures.C:74-90) and never
// fills CRP.CRP; the time-to-energy flux formula here is the textbook one:
//   A_e(R_i, theta_k) = sum_n exp(i E_e t_n) psi(R_i, r_s, theta_k, t_n) dt
//   B_e               = the same sum over d psi / dr at r_s
//   P(E_e) = Im( sum_omega sum_k w_k sum_i conj(A_e) B_e dR ) / (2 pi mu_r eta_sq(E_e))
// on the state with the leapfrogged V(dt/2) restored, t_n = (steps_n + 1) dt.
//
// Instead of a rank-1 update of the (n_E x surface) accumulators every step,
// the surface kernel only buffers the step's surface values and derivatives in
// a window of T steps; the window is then folded into A and B as one
// (n_E x T) . (T x surface) product, so the accumulators are read and written
// once per T steps.
#include "flux_surface.h"

namespace qd {

__global__ __launch_bounds__(kBlock) void k_flux_surface(StepArgs a) { flux_unit(a, blockIdx.x); }

void launch_flux(const StepArgs &a, int total_units, hipStream_t stream) {
  if (total_units <= 0) return;
  hipLaunchKernelGGL(k_flux_surface, dim3(total_units), dim3(kBlock), 0, stream, a);
}

// ph[e][q] = exp(i E_e t_q) dt for the steps of the window
__global__ __launch_bounds__(kBlock) void k_flux_phase(FluxWindow w) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= w.ne * w.n) return;
  const int e = idx / w.n, q = idx % w.n;
  const double t = (double)(w.s0 + q + 1) * w.dt;
  double sn, cs;
  sincos(w.E[e] * t, &sn, &cs);
  w.ph[(long)e * w.T + q] = make_double2(cs * w.dt, sn * w.dt);
}

// Tile: 64 surface points (one per lane) x 64 energies (16 per wave).
constexpr int kFluxEPW = 16;

__global__ __launch_bounds__(kBlock) void k_flux_window(FluxWindow w) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int s = blockIdx.x * 64 + lane;
  const int e0 = blockIdx.y * (4 * kFluxEPW) + wave * kFluxEPW;
  if (e0 >= w.ne) return;
  const int sc = s < w.S ? s : w.S - 1;   // clamped loads, predicated stores
  double2 accA[kFluxEPW], accB[kFluxEPW];
#pragma unroll
  for (int ee = 0; ee < kFluxEPW; ee++) {
    accA[ee] = make_double2(0.0, 0.0);
    accB[ee] = make_double2(0.0, 0.0);
  }
  for (int q = 0; q < w.n; q++) {
    const int slot = (w.s0 + q) % w.T;
    const double2 v = w.W[(long)slot * 2 * w.S + sc];
    const double2 d = w.W[((long)slot * 2 + 1) * w.S + sc];
#pragma unroll
    for (int ee = 0; ee < kFluxEPW; ee++) {
      const int e = min(e0 + ee, w.ne - 1);
      const double2 p = w.ph[(long)e * w.T + q];
      accA[ee] = cfma(p, v, accA[ee]);
      accB[ee] = cfma(p, d, accB[ee]);
    }
  }
  if (s >= w.S) return;
#pragma unroll
  for (int ee = 0; ee < kFluxEPW; ee++) {
    const int e = e0 + ee;
    if (e < w.ne) {
      const long o = (long)e * w.S + s;
      w.A[o] = cadd(w.A[o], accA[ee]);
      w.B[o] = cadd(w.B[o], accB[ee]);
    }
  }
}

void launch_flux_window(const FluxWindow &w, hipStream_t stream) {
  if (w.n <= 0) return;
  hipLaunchKernelGGL(k_flux_phase, dim3((w.ne * w.n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, w);
  const dim3 grid((w.S + 63) / 64, (w.ne + 4 * kFluxEPW - 1) / (4 * kFluxEPW));
  hipLaunchKernelGGL(k_flux_window, grid, dim3(kBlock), 0, stream, w);
}

// One workgroup per energy; fixed-order sums (deterministic).
__global__ __launch_bounds__(kBlock) void k_flux_finish(const double2 *__restrict__ A, const double2 *__restrict__ B,
                                                        const double *__restrict__ wk, const double *__restrict__ eta,
                                                        double *__restrict__ out, int S, int n1, int nt, double scale) {
  __shared__ double red[kBlock];
  const int e = blockIdx.x;
  const double2 *a = A + (long)e * S, *b = B + (long)e * S;
  double acc = 0.0;
  for (int s = threadIdx.x; s < S; s += kBlock) {
    const double2 x = a[s], y = b[s];
    acc = fma(wk[(s / n1) % nt], x.x * y.y - x.y * y.x, acc);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[e] = eta[e] > 0.0 ? red[0] * scale / eta[e] : 0.0;
}

void launch_flux_finish(const double2 *A, const double2 *B, const double *w, const double *eta, double *out, int S,
                        int n1, int nt, int ne, double scale, hipStream_t stream) {
  hipLaunchKernelGGL(k_flux_finish, dim3(ne), dim3(kBlock), 0, stream, A, B, w, eta, out, S, n1, nt, scale);
}

}  // namespace qd
