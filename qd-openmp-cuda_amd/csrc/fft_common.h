// fft_common.h - device helpers shared by the FFT passes: small DFT
// butterflies, the slab work-unit decoder and the per-point
// rotational . Coriolis . rotational mix.
#pragma once

#include "qd_complex.h"
#include "qd_internal.h"

namespace qd {

// Radix of the next stage when m = n / Ns points remain: 4s first, then 2,
// then 3s, then 5s.
__host__ __device__ constexpr int next_radix(int m) {
  return (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : 5;
}

__host__ __device__ constexpr int fft_stage_count(int n) {
  int s = 0;
  for (int ns = 1; ns < n; ns *= next_radix(n / ns)) s++;
  return s;
}

// Store of one complex value of an FFT pass's output, which the next pass
// reads once: streaming (nontemporal) when the job's arrays are larger than
// the Infinity Cache (stream, uniform), a plain store otherwise, so that the
// next pass finds it in the cache.
__device__ __forceinline__ void store_out(QD_G double2 *p, double2 v, int stream) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  if (stream) __builtin_nontemporal_store(d2{v.x, v.y}, (QD_G d2 *)p);
  else *p = v;
}

template <bool INV>
__device__ __forceinline__ double2 rot_i(double2 a) {   // multiply by -i (forward) / +i (inverse)
  return INV ? cmul_pi(a) : cmul_mi(a);
}

// a * exp(-+ i pi / 4): the radix-8 twiddle w^1 (forward -, inverse +)
template <bool INV>
__device__ __forceinline__ double2 rot_w8(double2 a) {
  constexpr double h = 0.70710678118654752440;
  return INV ? make_double2(h * (a.x - a.y), h * (a.x + a.y)) : make_double2(h * (a.x + a.y), h * (a.y - a.x));
}

// In-register DFT of R points, forward exp(-2 pi i qs/R) or inverse.
template <bool INV, int R>
__device__ __forceinline__ void dft_small(const double2 (&v)[R], double2 (&X)[R]) {
  if constexpr (R == 8) {
    // two 4-point DFTs of the even and odd points, combined with w^k
    const double2 ev[4] = {v[0], v[2], v[4], v[6]}, od[4] = {v[1], v[3], v[5], v[7]};
    double2 E[4], O[4];
    dft_small<INV, 4>(ev, E);
    dft_small<INV, 4>(od, O);
    const double2 o1 = rot_w8<INV>(O[1]), o2 = rot_i<INV>(O[2]), o3 = rot_i<INV>(rot_w8<INV>(O[3]));
    X[0] = cadd(E[0], O[0]);
    X[4] = csub(E[0], O[0]);
    X[1] = cadd(E[1], o1);
    X[5] = csub(E[1], o1);
    X[2] = cadd(E[2], o2);
    X[6] = csub(E[2], o2);
    X[3] = cadd(E[3], o3);
    X[7] = csub(E[3], o3);
  } else if constexpr (R == 2) {
    X[0] = cadd(v[0], v[1]);
    X[1] = csub(v[0], v[1]);
  } else if constexpr (R == 4) {
    const double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const double2 a2 = cadd(v[1], v[3]), a3 = rot_i<INV>(csub(v[1], v[3]));
    X[0] = cadd(a0, a2);
    X[2] = csub(a0, a2);
    X[1] = cadd(a1, a3);
    X[3] = csub(a1, a3);
  } else if constexpr (R == 3) {
    const double c3 = -0.5, s3 = 0.86602540378443864676;
    const double2 s = cadd(v[1], v[2]);
    const double2 d = rot_i<INV>(csub(v[1], v[2]));
    const double2 t = make_double2(v[0].x + c3 * s.x, v[0].y + c3 * s.y);
    X[0] = cadd(v[0], s);
    X[1] = make_double2(t.x + s3 * d.x, t.y + s3 * d.y);
    X[2] = make_double2(t.x - s3 * d.x, t.y - s3 * d.y);
  } else {
    static_assert(R == 5, "radix");
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

// Per-thread sums of the second FBR half step, per block: Coriolis energy of
// the pre-mix state, and of the output y the rotational energy, |y|^2 and
// conj(c0) y (c0: FBR coefficients of the initial state).
template <int NB, int NF>
struct FbrSums {
  static constexpr int F = NF;  // ce, re, norm [, acf re, acf im]
  double v[NB][F];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int f = 0; f < F; f++) v[b][f] = 0.0;
  }
};

// rotational . Coriolis . rotational at one grid point (i, j) of slab l for
// the nb blocks present at l:
//   y_b = rot2(l, j) sum_a V_l[a][b] phi(l, i, a) sum_b1 V_l[a][b1] x_b1
// = exp(-i dt/4 l(l+1)/(2 m r^2)) exp(-i dt/2 C_l/(2 m1 R^2)) exp(-i dt/4 ...) x
// with C_l = V_l^T diag(e) V_l; evolutionCUDAaux.cu:178-209 applies the same
// exponential as a dense sum over the eigenpairs.  With ENERGY, accumulates the
// Coriolis energy of x with the exact tridiagonal entries of C_l
// (evolutionCUDAaux.cu:211-241; the scalar rotational phase cancels in
// <x_b|C|x_b1>), the rotational energy l(l+1)(I1 + I2)|y|^2
// (omegawavepacket.cu:397-428), |y|^2 and, when c0 (this point of block 0 in
// the initial-state FBR array, blocks LM apart) is given, conj(c0) y.
// The last two equal the DVR norm and autocorrelation of the state after the
// backward Legendre transform: the Gauss-Legendre pair is an isometry between
// the FBR coefficients and the band-limited DVR values.
// phi points at the phases corPhi(l, i, a), a = 0.. (global table or an LDS copy).
template <bool ENERGY, int NB, int NF, class PhiPtr, class VPtr>
__device__ __forceinline__ void coriolis_point(const JobDev &jb, double2 (&x)[NB], int nb, int l, int i, int j,
                                               VPtr V, FbrSums<NB, NF> &sm,
                                               const QD_G double2 *c0, long LM, PhiPtr phi) {
  const int nO = jb.nO;
  if (ENERGY) {
    const double I1 = jb.I1[i];
#pragma unroll
    for (int b = 0; b < NB; b++) {
      if (b < nb) {
        const double2 c = jb.cmat[l * nO + b];
        double en = c.x * cnorm(x[b]);
        if (b + 1 < nb) en += 2.0 * c.y * cdot_re(x[b], x[b + 1]);
        sm.v[b][0] += en * I1;
      }
    }
  }
  double2 t[NB];
#pragma unroll
  for (int al = 0; al < NB; al++) {
    if (al < nb) {
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int b = 0; b < NB; b++)
        if (b < nb) {
          const double v = V[al * nO + b];
          acc.x = fma(v, x[b].x, acc.x);
          acc.y = fma(v, x[b].y, acc.y);
        }
      t[al] = cmul(phi[al], acc);
    }
  }
  const double2 r2 = jb.rot2[(long)l * jb.n2 + j];
  const double wrot = (double)l * (l + 1) * (jb.I1[i] + jb.I2[j]);
#pragma unroll
  for (int b = 0; b < NB; b++) {
    if (b < nb) {
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int al = 0; al < NB; al++)
        if (al < nb) {
          const double v = V[al * nO + b];
          acc.x = fma(v, t[al].x, acc.x);
          acc.y = fma(v, t[al].y, acc.y);
        }
      x[b] = cmul(r2, acc);
      if (ENERGY) {
        const double q = cnorm(x[b]);
        sm.v[b][1] += wrot * q;
        sm.v[b][2] += q;
        if constexpr (NF == 5) {
          const double2 z = c0[b * LM];
          sm.v[b][3] += z.x * x[b].x + z.y * x[b].y;
          sm.v[b][4] += z.x * x[b].y - z.y * x[b].x;
        }
      }
    }
  }
}

// Sum of x over the 64 lanes of the wave, written by lane 0 to *slot.
__device__ __forceinline__ void wave_sum_to(double x, double *slot) {
  x = wave_sum64(x);
  if ((threadIdx.x & 63) == 0) *slot = x;
}

// Sums the NF NB per-thread values over the block and stores them to
// part_cor[(5 b + f) * units + u]; red must hold 4 * NF * NB doubles.
template <int NB, int NF>
__device__ __forceinline__ void store_fbr_partials(const JobDev &jb, int u, const FbrSums<NB, NF> &sm, double *red) {
  constexpr int F = NF, NV = F * NB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    const double x = wave_sum64(sm.v[q / F][q % F]);
    if (lane == 0) red[wave * NV + q] = x;
  }
  __syncthreads();
  if (threadIdx.x < F * jb.nO) {
    const int q = threadIdx.x;   // q = F b + f, stored at field f of block b ([b][5][units])
    const double sum = (q < NV) ? (red[q] + red[NV + q]) + (red[2 * NV + q] + red[3 * NV + q]) : 0.0;
    jb.part_cor[(long)(5 * (q / F) + q % F) * jb.units[KK_FFT_RI] + u] = sum;
  }
}

}  // namespace qd
