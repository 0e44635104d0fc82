// fft_ct_impl.h - compile-time-length FFT passes for the common grid
// sizes (32, 48, 64, 96, 128, 192, 256 points along R or r).
//
// Same operators as kernels_fft.hip (see there for the reference
// counterparts), arranged as two passes (the kinetic split below).  With the
// length, the pencils per workgroup and the radix plan known at compile time,
// every index computation divides by constants and the stages unroll.  Each
// stage is in-place on one LDS tile: a thread reads the inputs of its
// butterflies into registers, the workgroup synchronises, and the outputs are
// written back, so a tile needs no ping-pong buffer.  In both passes the
// tiles of all omega blocks present at slab l stay in registers (the Coriolis
// mix couples blocks point by point) and only one block at a time passes
// through LDS.
#pragma once
#include "fft_common.h"

namespace qd {

// Stockham stages NS..N of an in-place transform of NPEN pencils; pencil p
// element e is A[p * PS + e * ES].  PF maps consecutive threads to
// consecutive pencils.  Entry and exit are synchronised.
template <bool INV, int N, int NS, int NPEN, int PS, int ES, bool PF>
__device__ __forceinline__ void stages_ip(double2 *A, const double2 *tw) {
  if constexpr (NS < N) {
    constexpr int R = next_radix(N / NS);
    constexpr int NB = N / R;
    constexpr int TOTAL = NPEN * NB;
    constexpr int MB = (TOTAL + kBlock - 1) / kBlock;
    constexpr int TWS = N / (NS * R);
    double2 X[MB][R];
#pragma unroll
    for (int m = 0; m < MB; m++) {
      const int t = threadIdx.x + m * kBlock;
      if (TOTAL % kBlock == 0 || t < TOTAL) {
        const int p = PF ? t % NPEN : t / NB;
        const int idx = PF ? t / NPEN : t % NB;
        double2 v[R];
#pragma unroll
        for (int q = 0; q < R; q++) v[q] = A[p * PS + (idx + q * NB) * ES];
        if constexpr (NS > 1) {
          const int k = idx % NS;
#pragma unroll
          for (int q = 1; q < R; q++) {
            double2 w = tw[q * k * TWS];
            if (INV) w.y = -w.y;
            v[q] = cmul(v[q], w);
          }
        }
        dft_small<INV, R>(v, X[m]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MB; m++) {
      const int t = threadIdx.x + m * kBlock;
      if (TOTAL % kBlock == 0 || t < TOTAL) {
        const int p = PF ? t % NPEN : t / NB;
        const int idx = PF ? t / NPEN : t % NB;
        const int d = (idx / NS) * NS * R + idx % NS;
#pragma unroll
        for (int q = 0; q < R; q++) A[p * PS + (d + q * NS) * ES] = X[m][q];
      }
    }
    __syncthreads();
    stages_ip<INV, N, NS * R, NPEN, PS, ES, PF>(A, tw);
  }
}

// Rows per workgroup of the R pass: 4 for n1 > 128; for shorter pencils 8,
// or 2 when more than 5 omega blocks are held in registers across the
// transform (measured: fewer rows lose more to the per-workgroup table
// staging than they gain in occupancy once n1 > 128).
__host__ __device__ constexpr int ct_rows(int n1, int nb) { return n1 > 128 ? 4 : (nb <= 5 ? 8 : 2); }
// Columns per workgroup of the r pass (all omega blocks of a slab in
// registers): 8 or 16 consecutive R points per r row keep 128-256 B
// contiguous per row; halved beyond 5 blocks or 128 r points.
__host__ __device__ constexpr int ct_cols(int n2, int nb) {
  return n2 <= 64 ? (nb <= 5 ? 16 : 8) : ((nb <= 5 && n2 <= 128) ? 8 : 4);
}

// Block-sums NB per-thread kinetic-energy values and stores block b's sum to
// part_ke[b * ke_stride + off] for every block of the job (zero where a block
// has no slab l).  red must hold 4 * NB doubles.
template <int NB>
__device__ __forceinline__ void store_ke_partials(const JobDev &jb, int off, const double (&ke)[NB], double *red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int b = 0; b < NB; b++) {
    double x = ke[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) red[wave * NB + b] = x;
  }
  __syncthreads();
  if (threadIdx.x < jb.nO) {
    const int b = threadIdx.x;
    const double sum = b < NB ? (red[b] + red[NB + b]) + (red[2 * NB + b] + red[3 * NB + b]) : 0.0;
    jb.part_ke[(long)b * jb.ke_stride + off] = sum;
  }
}

// ---------------------------------------------------------------------------
// The kinetic step as two one-dimensional round trips.  T = T_R + T_r with
// T_R acting on R only and T_r on r only, so they commute and
//   IFFT2 K FFT2 = [IFFT_R K_R FFT_R] [IFFT_r K_r FFT_r],
//   K_R = exp(-i dt k_R^2 / 2 m1), K_r = exp(-i dt k_r^2 / 2 m2),
// the same operator as the forward FFT, kinetic phase and scaled inverse FFT
// of cudaOpenMP2.cu:362-368.  Each round trip runs in LDS inside one pass, so
// the step needs two FFT passes over the FBR state instead of three:
//   R pass: first rotational-Coriolis-rotational half step, then
//           FFT_R . K_R . IFFT_R / n1 on every block;
//   r pass: FFT_r . K_r . IFFT_r / n2 on every block, then the second half
//           step with the Coriolis energy (of the state entering it), the
//           rotational energy, norm and autocorrelation.
// The kinetic energy of omegawavepacket.cu:364-395, sum |FFT2 psi|^2 T / M,
// splits the same way by Parseval: sum |FFT_R psi|^2 T_R / n1 in the R pass
// plus sum |FFT_r psi'|^2 T_r / n2 in the r pass (psi' = psi after the R
// round trip, which leaves <T_r> unchanged).
// ---------------------------------------------------------------------------
template <int NB, int N1>
__global__ __launch_bounds__(kBlock) void k_fft_Rk_ct(StepArgs a) {
  constexpr int P = ct_rows(N1, NB);
  constexpr int PN = P * N1;
  constexpr int PPT = (PN + kBlock - 1) / kBlock;
  __shared__ double2 A[PN];
  __shared__ double2 tw[N1];
  __shared__ double2 phs[N1 * NB];   // corPhi(l, i, a) of this slab, reused by the P rows
  __shared__ double2 kin[N1];
  __shared__ double kinE[N1];
  __shared__ double red[4 * NB];
  const int job = locate_job(a.unit_base, a.n_jobs, blockIdx.x);
  const JobDev &jb = a.jobs[job];
  const int u = blockIdx.x - a.unit_base[job];
  const int M = jb.M;
  const long LM = (long)jb.L * M;
  const int groups = jb.n2 / P;
  const int l = jb.omin + u / groups, g = u % groups;
  const int nb = min(jb.kmax, l) - jb.omin + 1;
  const QD_G double *V = jb.corV + (long)l * jb.nO * jb.nO;
  QD_G double2 *base = jb.fbr + (long)l * M + (long)g * PN;
  for (int e = threadIdx.x; e < N1; e += kBlock) {
    tw[e] = jb.twR[e];
    kin[e] = jb.kin1[e];
    kinE[e] = jb.kinE1[e];
  }
  {
    const QD_G double2 *cp = jb.corPhi + (long)l * N1 * jb.nO;
    for (int e = threadIdx.x; e < N1 * NB; e += kBlock) {
      const int i = e / NB, a2 = e % NB;
      const double2 v = cp[i * jb.nO + min(a2, jb.nO - 1)];
      phs[e] = a2 < jb.nO ? v : make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  // every global load of the slab tile in flight at once; loads are
  // unconditional from in-bounds addresses (block nb-1, last point), and
  // values of blocks >= nb and points >= PN are never used
  double2 x[NB][PPT];
#pragma unroll
  for (int b = 0; b < NB; b++)
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      const int e = threadIdx.x + k * kBlock;
      x[b][k] = base[(b < nb ? b : nb - 1) * LM + (PN % kBlock == 0 || e < PN ? e : PN - 1)];
    }
  FbrSums<NB, 3> sm;   // not accumulated: this half step forms no energies
  double ke[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) ke[b] = 0.0;
#pragma unroll
  for (int k = 0; k < PPT; k++) {
    const int e = threadIdx.x + k * kBlock;
    if (PN % kBlock == 0 || e < PN) {
      double2 xp[NB];
#pragma unroll
      for (int b = 0; b < NB; b++) xp[b] = x[b][k];
      coriolis_point<false, NB, 3>(jb, xp, nb, l, e % N1, g * P + e / N1, V, sm, nullptr, LM,
                                   (const double2 *)phs + (e % N1) * NB);
#pragma unroll
      for (int b = 0; b < NB; b++) x[b][k] = xp[b];
    }
  }
  const double s = 1.0 / (double)N1;
#pragma unroll
  for (int b = 0; b < NB; b++) {
    if (b < nb) {
#pragma unroll
      for (int k = 0; k < PPT; k++) {
        const int e = threadIdx.x + k * kBlock;
        if (PN % kBlock == 0 || e < PN) A[e] = x[b][k];
      }
      __syncthreads();
      stages_ip<false, N1, 1, P, N1, 1, false>(A, tw);
#pragma unroll
      for (int k = 0; k < PPT; k++) {
        const int e = threadIdx.x + k * kBlock;
        if (PN % kBlock == 0 || e < PN) {
          const double2 v = cmul(A[e], kin[e % N1]);
          ke[b] += cnorm(v) * kinE[e % N1];
          A[e] = v;
        }
      }
      __syncthreads();
      stages_ip<true, N1, 1, P, N1, 1, false>(A, tw);
#pragma unroll
      for (int k = 0; k < PPT; k++) {
        const int e = threadIdx.x + k * kBlock;
        if (PN % kBlock == 0 || e < PN) base[b * LM + e] = cscale(A[e], s);
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NB; b++) ke[b] *= s;
  store_ke_partials<NB>(jb, u, ke, red);
}

template <int NB, int N2, bool ACF>
__global__ __launch_bounds__(kBlock) void k_fft_rk_ct(StepArgs a) {
  constexpr int NF = ACF ? 5 : 3;
  constexpr int Q = ct_cols(N2, NB);
  constexpr int QN = Q * N2;
  constexpr int PPT = (QN + kBlock - 1) / kBlock;
  __shared__ double2 A[QN];
  __shared__ double2 tw[N2];
  __shared__ double2 phs[Q * NB];    // corPhi(l, i, a) for the Q columns i of this tile
  __shared__ double2 kin[N2];
  __shared__ double kinE[N2];
  __shared__ double red[4 * NF * NB];
  const int job = locate_job(a.unit_base, a.n_jobs, blockIdx.x);
  const JobDev &jb = a.jobs[job];
  const int u = blockIdx.x - a.unit_base[job];
  const int M = jb.M, n1 = jb.n1;
  const long LM = (long)jb.L * M;
  const int groups = n1 / Q;
  const int l = jb.omin + u / groups, g = u % groups;
  const int nb = min(jb.kmax, l) - jb.omin + 1;
  const QD_G double *V = jb.corV + (long)l * jb.nO * jb.nO;
  // tile element e: column i = g Q + e % Q, row j = e / Q (pencils along r)
  QD_G double2 *base = jb.fbr + (long)l * M + (long)g * Q;
  const QD_G double2 *c0 = ACF ? jb.c0 + (long)l * M + (long)g * Q : nullptr;
  for (int e = threadIdx.x; e < N2; e += kBlock) {
    tw[e] = jb.twr[e];
    kin[e] = jb.kin2[e];
    kinE[e] = jb.kinE2[e];
  }
  {
    const QD_G double2 *cp = jb.corPhi + ((long)l * n1 + (long)g * Q) * jb.nO;
    for (int e = threadIdx.x; e < Q * NB; e += kBlock) {
      const int i = e / NB, a2 = e % NB;
      const double2 v = cp[i * jb.nO + min(a2, jb.nO - 1)];
      phs[e] = a2 < jb.nO ? v : make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  auto goff = [&](int e) { return (long)(e / Q) * n1 + (e % Q); };
  double2 x[NB][PPT];
#pragma unroll
  for (int b = 0; b < NB; b++)
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      const int e = threadIdx.x + k * kBlock;
      x[b][k] = base[(b < nb ? b : nb - 1) * LM + goff(QN % kBlock == 0 || e < QN ? e : QN - 1)];
    }
  double ke[NB];
  const double s = 1.0 / (double)N2;
#pragma unroll
  for (int b = 0; b < NB; b++) {
    ke[b] = 0.0;
    if (b < nb) {
#pragma unroll
      for (int k = 0; k < PPT; k++) {
        const int e = threadIdx.x + k * kBlock;
        if (QN % kBlock == 0 || e < QN) A[e] = x[b][k];
      }
      __syncthreads();
      stages_ip<false, N2, 1, Q, 1, Q, true>(A, tw);
#pragma unroll
      for (int k = 0; k < PPT; k++) {
        const int e = threadIdx.x + k * kBlock;
        if (QN % kBlock == 0 || e < QN) {
          const double2 v = cmul(A[e], kin[e / Q]);
          ke[b] += cnorm(v) * kinE[e / Q];
          A[e] = v;
        }
      }
      __syncthreads();
      stages_ip<true, N2, 1, Q, 1, Q, true>(A, tw);
#pragma unroll
      for (int k = 0; k < PPT; k++) {
        const int e = threadIdx.x + k * kBlock;
        x[b][k] = (QN % kBlock == 0 || e < QN) ? cscale(A[e], s) : make_double2(0.0, 0.0);
      }
      ke[b] *= s;
    } else {
#pragma unroll
      for (int k = 0; k < PPT; k++) x[b][k] = make_double2(0.0, 0.0);
    }
  }
  FbrSums<NB, NF> sm;
  sm.zero();
#pragma unroll
  for (int k = 0; k < PPT; k++) {
    const int e = threadIdx.x + k * kBlock;
    if (QN % kBlock == 0 || e < QN) {
      double2 xp[NB];
#pragma unroll
      for (int b = 0; b < NB; b++) xp[b] = x[b][k];
      coriolis_point<true, NB, NF>(jb, xp, nb, l, g * Q + e % Q, e / Q, V, sm, ACF ? c0 + goff(e) : nullptr, LM,
                                   (const double2 *)phs + (e % Q) * NB);
#pragma unroll
      for (int b = 0; b < NB; b++)
        if (b < nb) base[b * LM + goff(e)] = xp[b];
    }
  }
  store_fbr_partials<NB, NF>(jb, u, sm, red);
  __syncthreads();
  store_ke_partials<NB>(jb, jb.ke_off_r + u, ke, red);
}

// ---------------------------------------------------------------------------
// launchers for one compile-time length (instantiated in kernels_fft_ct*.hip)
// ---------------------------------------------------------------------------
template <int N1>
void launch_Rk_nb(const StepArgs &a, int units, int nb_max, hipStream_t s) {
  const dim3 g(units), t(kBlock);
  if (nb_max <= 1) hipLaunchKernelGGL((k_fft_Rk_ct<1, N1>), g, t, 0, s, a);
  else if (nb_max <= 2) hipLaunchKernelGGL((k_fft_Rk_ct<2, N1>), g, t, 0, s, a);
  else if (nb_max <= 3) hipLaunchKernelGGL((k_fft_Rk_ct<3, N1>), g, t, 0, s, a);
  else if (nb_max <= 5) hipLaunchKernelGGL((k_fft_Rk_ct<5, N1>), g, t, 0, s, a);
  else if (nb_max <= 7) hipLaunchKernelGGL((k_fft_Rk_ct<7, N1>), g, t, 0, s, a);
  else if (nb_max <= 9) hipLaunchKernelGGL((k_fft_Rk_ct<9, N1>), g, t, 0, s, a);
  else if (nb_max <= 16) hipLaunchKernelGGL((k_fft_Rk_ct<16, N1>), g, t, 0, s, a);
  else hipLaunchKernelGGL((k_fft_Rk_ct<24, N1>), g, t, 0, s, a);
}

template <int N2, bool ACF>
void launch_rk_nb(const StepArgs &a, int units, int nb_max, hipStream_t s) {
  const dim3 g(units), t(kBlock);
  if (nb_max <= 1) hipLaunchKernelGGL((k_fft_rk_ct<1, N2, ACF>), g, t, 0, s, a);
  else if (nb_max <= 2) hipLaunchKernelGGL((k_fft_rk_ct<2, N2, ACF>), g, t, 0, s, a);
  else if (nb_max <= 3) hipLaunchKernelGGL((k_fft_rk_ct<3, N2, ACF>), g, t, 0, s, a);
  else if (nb_max <= 5) hipLaunchKernelGGL((k_fft_rk_ct<5, N2, ACF>), g, t, 0, s, a);
  else if (nb_max <= 7) hipLaunchKernelGGL((k_fft_rk_ct<7, N2, ACF>), g, t, 0, s, a);
  else if (nb_max <= 9) hipLaunchKernelGGL((k_fft_rk_ct<9, N2, ACF>), g, t, 0, s, a);
  else if (nb_max <= 16) hipLaunchKernelGGL((k_fft_rk_ct<16, N2, ACF>), g, t, 0, s, a);
  else hipLaunchKernelGGL((k_fft_rk_ct<24, N2, ACF>), g, t, 0, s, a);
}

#define QD_FFT_CT_DEFINE(N)                                                                    \
  void launch_fft_Rk_ct_##N(const StepArgs &a, int units, int nb_max, hipStream_t s) {         \
    launch_Rk_nb<N>(a, units, nb_max, s);                                                      \
  }                                                                                            \
  void launch_fft_rk_ct_##N(const StepArgs &a, int units, int acf, int nb_max, hipStream_t s) { \
    if (acf) launch_rk_nb<N, true>(a, units, nb_max, s);                                       \
    else launch_rk_nb<N, false>(a, units, nb_max, s);                                          \
  }

}  // namespace qd
