/*
 * oracle.cpp - FP64 CPU restatement of the reference's split-operator step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP
 * propagator: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links it.
 *
 * It follows the reference file by file (citations are into the reference
 * sources):
 *   step order ........... cudaOpenMP2.cu:338-387 (test_multiple_cards)
 *   potential phase ...... evolutionCUDAaux.cu:5-11, omegawavepacket.cu:314-328
 *   absorber ............. evolutionCUDAaux.cu:243-252, omegawavepacket.cu:491-501
 *   Legendre transforms .. omegawavepacket.cu:134-240 (P, wP tables, ZGEMMs)
 *   rotational phase ..... evolutionCUDAaux.cu:33-54, cudaMath.h:150-157
 *   Coriolis ............. CoriolisMatrix.C:23-99, cudaOpenMP2.cu:158-209,
 *                          evolutionCUDAaux.cu:178-241, omegawavepacket.cu:430-489, :503-564,
 *                          wavepackets1device.cu:485-530
 *   2-D FFT .............. wavepackets1device.cu:186-201, omegawavepacket.cu:242-269
 *   kinetic phase ........ evolutionCUDAaux.cu:13-31, cudaMath.h:117-130
 *   observables .......... omegawavepacket.cu:76-132, :364-428; cudaOpenMP2.cu:425-445
 *   copy-back / steps .... cudaOpenMP2.cu:380-386, omegawavepacket.cu:271-280
 * Its own primitives replace the vendor libraries: a recursive mixed-radix
 * FFT (cuFFT Z2Z), a cyclic Jacobi eigensolver (LAPACK dstev), plain loops
 * (cuBLAS ZGEMM / ZDOTC / ZDSCAL).
 *
 * Extensions that the reference does not have (K_max truncation of the
 * Coriolis matrices, autocorrelation, flux / reaction probabilities) follow
 * the definitions written in include/qd.h.
 *
 * Pinning (tests/test_oracle.py): FFT against brute-force DFT and numpy.fft
 * golden vectors; eigensolver against numpy eigh golden vectors and
 * V diag(e) V^T = C, V^T V = I; Legendre tables against scipy and
 * quadrature orthonormality; the whole step against free-particle and
 * rigid-rotor closed forms, time reversibility, norm and energy conservation,
 * the n = 2 Coriolis closed form and the J = 0 identity; flux against the
 * Eckart barrier transmission; the autocorrelation against the free-particle
 * closed form <psi_0|IFFT2(exp(-i t T) FFT2 psi_0)>, against the phase
 * exp(-i V0 t) that a constant potential V0 adds, and across a resume.
 */
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "qd.h"

typedef std::complex<double> cplx;

namespace {

// ---------------------------------------------------------------------------
// FFT: recursive decimation-in-time, any length, smallest-prime-factor first.
// ---------------------------------------------------------------------------
struct FFTPlan {
  int n = 0;
  std::vector<cplx> W;  // W[e] = exp(-2 pi i e / n)
  explicit FFTPlan(int n_) : n(n_), W(n_) {
    for (int e = 0; e < n; e++) {
      long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)e / (long double)n;
      W[e] = cplx((double)cosl(a), (double)sinl(a));
    }
  }
};

static int smallest_factor(int n) {
  for (int p = 2; p * p <= n; p++)
    if (n % p == 0) return p;
  return n;
}

// y[0..n) = DFT_n of x[0], x[xs], ..., sign -1 forward / +1 inverse (unnormalised)
static void fft_rec(const cplx *x, long xs, cplx *y, int n, const FFTPlan &pl, int sign) {
  if (n == 1) { y[0] = x[0]; return; }
  const int p = smallest_factor(n), m = n / p;
  for (int q = 0; q < p; q++) fft_rec(x + q * xs, xs * p, y + q * m, m, pl, sign);
  const int N = pl.n, sn = N / n, sp = N / p;
  cplx t[64], o[64];
  std::vector<cplx> tb, ob;
  cplx *tt = t, *oo = o;
  if (p > 64) { tb.resize(p); ob.resize(p); tt = tb.data(); oo = ob.data(); }
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      cplx w = pl.W[((long)q * k * sn) % N];
      if (sign > 0) w = std::conj(w);
      tt[q] = y[q * m + k] * w;
    }
    for (int s = 0; s < p; s++) {
      cplx acc = 0;
      for (int q = 0; q < p; q++) {
        cplx w = pl.W[((long)q * s * sp) % N];
        if (sign > 0) w = std::conj(w);
        acc += tt[q] * w;
      }
      oo[s] = acc;
    }
    for (int s = 0; s < p; s++) y[k + s * m] = oo[s];
  }
}

// in-place strided transform of n elements
static void fft_strided(cplx *a, long stride, int n, const FFTPlan &pl, int sign, cplx *buf_in, cplx *buf_out) {
  for (int i = 0; i < n; i++) buf_in[i] = a[i * stride];
  fft_rec(buf_in, 1, buf_out, n, pl, sign);
  for (int i = 0; i < n; i++) a[i * stride] = buf_out[i];
}

// ---------------------------------------------------------------------------
// Symmetric eigensolver: cyclic Jacobi.  a is n x n column-major (destroyed).
// On return e[alpha] are eigenvalues, v[alpha*n + i] component i of vector alpha
// (the column-major layout dstev writes, CoriolisMatrix.C:89-94).
// ---------------------------------------------------------------------------
static void jacobi_eig(int n, double *a, double *e, double *v) {
  for (int i = 0; i < n * n; i++) v[i] = 0.0;
  for (int i = 0; i < n; i++) v[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) off += a[q * n + p] * a[q * n + p];
    if (off == 0.0) break;
    for (int p = 0; p < n; p++) {
      for (int q = p + 1; q < n; q++) {
        const double apq = a[q * n + p];
        if (apq == 0.0) continue;
        const double app = a[p * n + p], aqq = a[q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {  // columns p, q of A
          const double akp = a[p * n + k], akq = a[q * n + k];
          a[p * n + k] = c * akp - s * akq;
          a[q * n + k] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {  // rows p, q of A
          const double apk = a[k * n + p], aqk = a[k * n + q];
          a[k * n + p] = c * apk - s * aqk;
          a[k * n + q] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {  // eigenvectors: V <- V J
          const double vkp = v[p * n + k], vkq = v[q * n + k];
          v[p * n + k] = c * vkp - s * vkq;
          v[q * n + k] = s * vkp + c * vkq;
        }
      }
    }
  }
  for (int i = 0; i < n; i++) e[i] = a[i * n + i];
}

// lambda(J, Omega) = sqrt(J(J+1) - Omega(Omega+1)) for J >= Omega, else 0 (CoriolisMatrix.C:23-33)
static double lambda_plus(int J, int om) {
  return J >= om ? std::sqrt(J * (J + 1.0) - om * (om + 1.0)) : 0.0;
}

// Dense tridiagonal Coriolis matrix C_l over omega = omega_min..omega_min+n-1
// (CoriolisMatrix.C:63-87): diag J(J+1) - 2 omega^2,
// off-diag -lambda(J,omega) lambda(l,omega) sqrt(1 + delta_{omega,0}).
static void coriolis_dense(int J, int l, int omega_min, int n, double *c) {
  for (int i = 0; i < n * n; i++) c[i] = 0.0;
  for (int k = 0; k < n; k++) {
    const int om = k + omega_min;
    c[k * n + k] = J * (J + 1) - 2 * om * om;
  }
  for (int k = 0; k < n - 1; k++) {
    const int om = k + omega_min;
    const double s = -lambda_plus(J, om) * lambda_plus(l, om) * std::sqrt(1.0 + (om == 0 ? 1.0 : 0.0));
    c[(k + 1) * n + k] = s;
    c[k * n + (k + 1)] = s;
  }
}

struct CorMat {
  int l = -1, n = 0;
  std::vector<double> e, v;
};

struct Job {
  int n1, n2, nt, M, L, lmax, nO, omin, J, p, kmax;
  double dt, dr1, dr2, left1, left2, m1, m2;
  std::vector<int> om;                     // omegas
  std::vector<std::vector<cplx>> psi;      // [b] G
  std::vector<std::vector<cplx>> fbr;      // [b] L*M
  std::vector<std::vector<cplx>> scratch;  // [b] L*M (Coriolis accumulator)
  std::vector<std::vector<double>> P;      // [b] (nt x Nl) column-major P(k, l-omega)
  std::vector<std::vector<double>> wP;     // [b] w_k P(k, l-omega)
  std::vector<double> pot, w, kin1, kin2, I1, I2, d1, d2;
  bool dump = false;
  std::vector<CorMat> cor;                 // [l]
  std::vector<cplx> psi0;                  // state at the start time t_0, all blocks (autocorrelation)
};

static double kinetic_entry(int i, int n, double xl, double mass) {
  // cudaMath.h:117-130
  const double two_pi_xl = 2 * M_PI / xl;
  const double k = (i <= n / 2) ? two_pi_xl * i : two_pi_xl * (-n + i);
  return k * k / (mass + mass);
}

static cplx expi(double x) { return cplx(std::cos(x), std::sin(x)); }

}  // namespace

// ---------------------------------------------------------------------------
// exported primitives (for pinning the oracle itself)
// ---------------------------------------------------------------------------
extern "C" void oracle_fft1d(int n, int sign, double *data) {
  FFTPlan pl(n);
  std::vector<cplx> in(n), out(n);
  cplx *a = reinterpret_cast<cplx *>(data);
  fft_strided(a, 1, n, pl, sign, in.data(), out.data());
}

extern "C" void oracle_sym_eig(int n, const double *a_in, double *e, double *v) {
  std::vector<double> a(a_in, a_in + n * n);
  jacobi_eig(n, a.data(), e, v);
}

extern "C" void oracle_coriolis_dense(int J, int l, int omega_min, int n, double *c) {
  coriolis_dense(J, l, omega_min, n, c);
}

// Coriolis exponential of the oracle's step: out[b*n + b1] (complex, interleaved)
// = sum_a exp(-i tau e_a / I) v_ab v_ab1 (evolutionCUDAaux.cu:185-198) with
// (e, v) the eigenpairs of the dense C_l over omega_min..omega_min+n-1.
extern "C" void oracle_coriolis_exp(int J, int l, int omega_min, int n, double tau, double I, double *out) {
  std::vector<double> a(n * n), e(n), v(n * n);
  coriolis_dense(J, l, omega_min, n, a.data());
  jacobi_eig(n, a.data(), e.data(), v.data());
  for (int b = 0; b < n; b++)
    for (int b1 = 0; b1 < n; b1++) {
      cplx acc = 0.0;
      for (int al = 0; al < n; al++) acc += expi(-tau / I * e[al]) * v[al * n + b] * v[al * n + b1];
      out[2 * (b * n + b1)] = acc.real();
      out[2 * (b * n + b1) + 1] = acc.imag();
    }
}

// ---------------------------------------------------------------------------
// the propagator
// ---------------------------------------------------------------------------
static int setup_job(const qd_inputs *in, Job &jb, char *err) {
  const qd_omega_states &os = in->omega_states;
  jb.n1 = in->r1.n; jb.n2 = in->r2.n; jb.nt = in->theta.n;
  jb.M = jb.n1 * jb.n2;
  jb.lmax = os.l_max; jb.L = os.l_max + 1;
  jb.nO = os.n_omegas; jb.J = os.J; jb.p = os.parity;
  jb.dt = in->time.time_step;
  jb.dr1 = in->r1.dr; jb.dr2 = in->r2.dr; jb.left1 = in->r1.left; jb.left2 = in->r2.left;
  jb.m1 = in->r1.mass; jb.m2 = in->r2.mass;
  if (jb.nt <= jb.lmax + 1) { snprintf(err, 256, "ntheta must exceed l_max+1"); return 1; }
  if (jb.n1 % 2 || jb.n2 % 2) { snprintf(err, 256, "n1 and n2 must be even"); return 1; }
  jb.om.assign(os.omegas, os.omegas + jb.nO);
  jb.omin = jb.om[0];
  const int kcap = std::min(jb.J, jb.lmax);
  jb.kmax = in->k_max >= 0 ? std::min(kcap, (int)in->k_max) : kcap;
  const long G = (long)jb.M * jb.nt;
  jb.psi.resize(jb.nO); jb.fbr.resize(jb.nO); jb.scratch.resize(jb.nO);
  jb.P.resize(jb.nO); jb.wP.resize(jb.nO);
  jb.w.assign(in->theta.w, in->theta.w + jb.nt);
  const int nmax = os.legendre_n_max;
  for (int b = 0; b < jb.nO; b++) {
    const cplx *src = reinterpret_cast<const cplx *>(os.wave_packets) + b * G;
    jb.psi[b].assign(src, src + G);
    jb.fbr[b].assign((long)jb.L * jb.M, cplx(0, 0));
    jb.scratch[b].assign((long)jb.L * jb.M, cplx(0, 0));
    const int nl = jb.lmax - jb.om[b] + 1;
    jb.P[b].resize((long)jb.nt * nl);
    jb.wP[b].resize((long)jb.nt * nl);
    const double *pb = os.associated_legendres + (long)b * jb.nt * nmax;
    for (int c = 0; c < nl; c++)
      for (int k = 0; k < jb.nt; k++) {
        jb.P[b][(long)c * jb.nt + k] = pb[(long)c * jb.nt + k];
        jb.wP[b][(long)c * jb.nt + k] = jb.w[k] * pb[(long)c * jb.nt + k];
      }
  }
  jb.pot.assign(in->pot, in->pot + G);
  // psi(t_0): the caller's array on a fresh start (steps == 0); on a resume
  // the array is a stored state, which lacks the trailing V(dt/2) of the
  // leapfrogged potential (cudaOpenMP2.cu:338-387), so it is restored here
  jb.psi0.resize(G * jb.nO);
  const bool resumed = *in->time.steps != 0;
  for (int b = 0; b < jb.nO; b++)
    for (long idx = 0; idx < G; idx++)
      jb.psi0[b * G + idx] = resumed ? jb.psi[b][idx] * expi(-0.5 * jb.dt * jb.pot[idx]) : jb.psi[b][idx];
  jb.kin1.resize(jb.n1); jb.kin2.resize(jb.n2); jb.I1.resize(jb.n1); jb.I2.resize(jb.n2);
  for (int i = 0; i < jb.n1; i++) {
    jb.kin1[i] = kinetic_entry(i, jb.n1, jb.n1 * jb.dr1, jb.m1);
    const double r = jb.left1 + i * jb.dr1;
    jb.I1[i] = 1.0 / (2 * jb.m1 * r * r);   // cudaMath.h:150-157
  }
  for (int j = 0; j < jb.n2; j++) {
    jb.kin2[j] = kinetic_entry(j, jb.n2, jb.n2 * jb.dr2, jb.m2);
    const double r = jb.left2 + j * jb.dr2;
    jb.I2[j] = 1.0 / (2 * jb.m2 * r * r);
  }
  jb.dump = in->dump1.dump && in->dump2.dump;
  if (jb.dump) {
    jb.d1.assign(in->dump1.dump, in->dump1.dump + jb.n1);
    jb.d2.assign(in->dump2.dump, in->dump2.dump + jb.n2);
  }
  // Coriolis pack for l = omegas[0]..l_max (cudaOpenMP2.cu:158-209)
  jb.cor.assign(jb.L, CorMat());
  for (int l = jb.omin; l <= jb.lmax; l++) {
    CorMat &c = jb.cor[l];
    c.l = l;
    c.n = std::min(jb.kmax, l) - jb.omin + 1;
    std::vector<double> a(c.n * c.n);
    coriolis_dense(jb.J, l, jb.omin, c.n, a.data());
    c.e.resize(c.n); c.v.resize(c.n * c.n);
    jacobi_eig(c.n, a.data(), c.e.data(), c.v.data());
  }
  return 0;
}

static void phase_dvr(Job &jb, double tau) {   // psi *= exp(-i tau V)
  const long G = (long)jb.M * jb.nt;
  for (int b = 0; b < jb.nO; b++) {
    cplx *ps = jb.psi[b].data();
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < G; idx++) ps[idx] *= expi(-tau * jb.pot[idx]);
  }
}

static void dump_dvr(Job &jb) {
  for (int b = 0; b < jb.nO; b++) {
    cplx *ps = jb.psi[b].data();
#pragma omp parallel for schedule(static)
    for (int kj = 0; kj < jb.nt * jb.n2; kj++) {
      const int j = kj % jb.n2;
      for (int i = 0; i < jb.n1; i++) ps[(long)kj * jb.n1 + i] *= jb.d1[i] * jb.d2[j];
    }
  }
}

static void legendre_forward(Job &jb) {
  const int M = jb.M;
  for (int b = 0; b < jb.nO; b++) {
    const int om = jb.om[b], nl = jb.lmax - om + 1;
#pragma omp parallel for schedule(dynamic)
    for (int c = 0; c < nl; c++) {
      cplx *out = jb.fbr[b].data() + (long)(om + c) * M;
      for (int m = 0; m < M; m++) out[m] = 0.0;
      for (int k = 0; k < jb.nt; k++) {
        const double wp = jb.wP[b][(long)c * jb.nt + k];
        const cplx *src = jb.psi[b].data() + (long)k * M;
        for (int m = 0; m < M; m++) out[m] += src[m] * wp;
      }
    }
  }
}

static void legendre_backward(Job &jb) {
  const int M = jb.M;
  for (int b = 0; b < jb.nO; b++) {
    const int om = jb.om[b], nl = jb.lmax - om + 1;
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < jb.nt; k++) {
      cplx *out = jb.psi[b].data() + (long)k * M;
      for (int m = 0; m < M; m++) out[m] = 0.0;
      for (int c = 0; c < nl; c++) {
        const double p = jb.P[b][(long)c * jb.nt + k];
        const cplx *src = jb.fbr[b].data() + (long)(om + c) * M;
        for (int m = 0; m < M; m++) out[m] += src[m] * p;
      }
    }
  }
}

static void rotational(Job &jb, double tau) {
  const int M = jb.M;
  for (int b = 0; b < jb.nO; b++) {
    const int om = jb.om[b];
#pragma omp parallel for schedule(static)
    for (int l = om; l <= jb.lmax; l++) {
      cplx *s = jb.fbr[b].data() + (long)l * M;
      for (int j = 0; j < jb.n2; j++)
        for (int i = 0; i < jb.n1; i++)
          s[j * jb.n1 + i] *= expi(-tau * l * (l + 1) * (jb.I1[i] + jb.I2[j]));
    }
  }
}

static void kinetic(Job &jb, double tau) {
  const int M = jb.M;
  for (int b = 0; b < jb.nO; b++) {
    const int om = jb.om[b];
#pragma omp parallel for schedule(static)
    for (int l = om; l <= jb.lmax; l++) {
      cplx *s = jb.fbr[b].data() + (long)l * M;
      for (int j = 0; j < jb.n2; j++)
        for (int i = 0; i < jb.n1; i++)
          s[j * jb.n1 + i] *= expi(-tau * (jb.kin1[i] + jb.kin2[j]));
    }
  }
}

static void fft2d(Job &jb, int sign) {
  FFTPlan p1(jb.n1), p2(jb.n2);
  const int M = jb.M;
  for (int b = 0; b < jb.nO; b++) {
    const int om = jb.om[b];
#pragma omp parallel
    {
      std::vector<cplx> bi(std::max(jb.n1, jb.n2)), bo(std::max(jb.n1, jb.n2));
#pragma omp for schedule(static)
      for (int l = om; l <= jb.lmax; l++) {
        cplx *s = jb.fbr[b].data() + (long)l * M;
        for (int j = 0; j < jb.n2; j++) fft_strided(s + (long)j * jb.n1, 1, jb.n1, p1, sign, bi.data(), bo.data());
        for (int i = 0; i < jb.n1; i++) fft_strided(s + i, jb.n1, jb.n2, p2, sign, bi.data(), bo.data());
      }
    }
  }
}

static void scale_fbr(Job &jb) {
  const double s = 1.0 / (jb.n1 * jb.n2);
  for (int b = 0; b < jb.nO; b++) {
    const long lo = (long)jb.om[b] * jb.M, hi = (long)jb.L * jb.M;
    cplx *f = jb.fbr[b].data();
#pragma omp parallel for schedule(static)
    for (long idx = lo; idx < hi; idx++) f[idx] *= s;
  }
}

// one Coriolis half step; energies written to core[b] when non-null
static void coriolis(Job &jb, double tau, double *core) {
  const int M = jb.M, n1 = jb.n1, n2 = jb.n2;
  if (core)
    for (int b = 0; b < jb.nO; b++) core[b] = 0.0;
  // energy on the pre-update state (wavepackets1device.cu:515-530)
  if (core) {
    for (int b = 0; b < jb.nO; b++) {
      for (int b1 = b; b1 < jb.nO; b1++) {
        const int o = jb.om[b], o1 = jb.om[b1];
        if (!(o == o1 || o + 1 == o1)) continue;
        double sum = 0.0;
#pragma omp parallel for reduction(+ : sum) schedule(static)
        for (int l = std::max(o1, jb.omin); l <= jb.lmax; l++) {
          const CorMat &cm = jb.cor[l];
          if (o - jb.omin >= cm.n || o1 - jb.omin >= cm.n) continue;
          const cplx *x = jb.fbr[b].data() + (long)l * M;
          const cplx *y = jb.fbr[b1].data() + (long)l * M;
          cplx dot = 0.0;
          for (int i = 0; i < n1; i++) {
            double c = 0.0;   // evolutionCUDAaux.cu:219-228
            for (int a = 0; a < cm.n; a++) c += cm.e[a] * cm.v[a * cm.n + (o - jb.omin)] * cm.v[a * cm.n + (o1 - jb.omin)];
            const double R = jb.left1 + i * jb.dr1;
            c /= 2 * jb.m1 * R * R;
            for (int j = 0; j < n2; j++) dot += std::conj(x[j * n1 + i]) * (c * y[j * n1 + i]);
          }
          if (o != o1) dot *= 2.0;
          sum += dot.real() * jb.dr1 * jb.dr2;
        }
        core[b] += sum;
      }
    }
  }
  // accumulate the exponential mix (evolutionCUDAaux.cu:178-209)
  for (int b = 0; b < jb.nO; b++) {
    const int o = jb.om[b];
#pragma omp parallel
    {
      std::vector<cplx> expC(n1);
#pragma omp for schedule(static)
      for (int l = o; l <= jb.lmax; l++) {
        const CorMat &cm = jb.cor[l];
        cplx *out = jb.scratch[b].data() + (long)l * M;
        for (int m = 0; m < M; m++) out[m] = 0.0;
        for (int b1 = 0; b1 < jb.nO; b1++) {
          const int o1 = jb.om[b1];
          if (o1 > l || o - jb.omin >= cm.n || o1 - jb.omin >= cm.n) continue;
          for (int i = 0; i < n1; i++) {
            const double R = jb.left1 + i * jb.dr1;
            const double I = 2 * jb.m1 * R * R;
            cplx acc = 0.0;
            for (int a = 0; a < cm.n; a++)
              acc += expi(-tau / I * cm.e[a]) * cm.v[a * cm.n + (o - jb.omin)] * cm.v[a * cm.n + (o1 - jb.omin)];
            expC[i] = acc;
          }
          const cplx *src = jb.fbr[b1].data() + (long)l * M;
          for (int j = 0; j < n2; j++)
            for (int i = 0; i < n1; i++) out[j * n1 + i] += expC[i] * src[j * n1 + i];
        }
      }
    }
  }
  for (int b = 0; b < jb.nO; b++) {   // omegawavepacket.cu:479-489
    const long lo = (long)jb.om[b] * M, hi = (long)jb.L * M;
    std::copy(jb.scratch[b].begin() + lo, jb.scratch[b].begin() + hi, jb.fbr[b].begin() + lo);
  }
}

static void kinetic_energy(Job &jb, double *ke) {
  for (int b = 0; b < jb.nO; b++) {
    double sum = 0.0;
#pragma omp parallel for reduction(+ : sum) schedule(static)
    for (int l = jb.om[b]; l <= jb.lmax; l++) {
      const cplx *s = jb.fbr[b].data() + (long)l * jb.M;
      double t = 0.0;
      for (int j = 0; j < jb.n2; j++)
        for (int i = 0; i < jb.n1; i++) t += std::norm(s[j * jb.n1 + i]) * (jb.kin1[i] + jb.kin2[j]);
      sum += t;
    }
    ke[b] = sum * jb.dr1 * jb.dr2 / (jb.n1 * jb.n2);
  }
}

static void rotational_energy(Job &jb, double *re) {
  for (int b = 0; b < jb.nO; b++) {
    double sum = 0.0;
#pragma omp parallel for reduction(+ : sum) schedule(static)
    for (int l = jb.om[b]; l <= jb.lmax; l++) {
      const cplx *s = jb.fbr[b].data() + (long)l * jb.M;
      double t = 0.0;
      for (int j = 0; j < jb.n2; j++)
        for (int i = 0; i < jb.n1; i++) t += std::norm(s[j * jb.n1 + i]) * (jb.I1[i] + jb.I2[j]);
      sum += l * (l + 1) * t;
    }
    re[b] = sum * jb.dr1 * jb.dr2;
  }
}

static void dvr_observables(Job &jb, double *norm, double *pe) {
  const int M = jb.M;
  for (int b = 0; b < jb.nO; b++) {
    double sn = 0.0, sp = 0.0;
#pragma omp parallel for reduction(+ : sn, sp) schedule(static)
    for (int k = 0; k < jb.nt; k++) {
      const cplx *s = jb.psi[b].data() + (long)k * M;
      const double *v = jb.pot.data() + (long)k * M;
      double a = 0.0, c = 0.0;
      for (int m = 0; m < M; m++) { const double q = std::norm(s[m]); a += q; c += q * v[m]; }
      sn += jb.w[k] * a;
      sp += jb.w[k] * c;
    }
    norm[b] = sn * jb.dr1 * jb.dr2;
    pe[b] = sp * jb.dr1 * jb.dr2;
  }
}

// C(t_n) = <psi(t_0)|psi(t_n)> = sum_b sum_k w_k sum_ij conj(psi0) psi dr1 dr2, with
// psi(t_n) the stored state with the missing V(dt/2) restored (as for the flux).
static cplx autocorrelation(Job &jb) {
  const int M = jb.M;
  const long G = (long)M * jb.nt;
  double sr = 0.0, si = 0.0;
  for (int b = 0; b < jb.nO; b++) {
#pragma omp parallel for reduction(+ : sr, si) schedule(static)
    for (int k = 0; k < jb.nt; k++) {
      cplx acc = 0.0;
      const cplx *s = jb.psi[b].data() + (long)k * M;
      const cplx *s0 = jb.psi0.data() + b * G + (long)k * M;
      const double *v = jb.pot.data() + (long)k * M;
      for (int m = 0; m < M; m++) acc += std::conj(s0[m]) * (s[m] * expi(-0.5 * jb.dt * v[m]));
      sr += jb.w[k] * acc.real();
      si += jb.w[k] * acc.imag();
    }
  }
  return cplx(sr, si) * (jb.dr1 * jb.dr2);
}

// Flux accumulation at r_s: A_e += exp(i E_e t) psi dt, B_e += exp(i E_e t) dpsi/dr dt,
// on the state with the missing V(dt/2) restored.
struct Flux {
  int ns = 0, ng = 0, ne = 0;
  std::vector<double> coef, E, eta;
  std::vector<cplx> A, B;   // [e][b][k][i]
};

static void flux_accumulate(Job &jb, Flux &fx, double t) {
  const int n1 = jb.n1, M = jb.M, half = fx.ng / 2;
  const long S = (long)jb.nO * jb.nt * n1;
  std::vector<cplx> val(S), der(S);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < S; s++) {
    const int i = s % n1;
    const int k = (s / n1) % jb.nt;
    const int b = s / ((long)n1 * jb.nt);
    cplx d = 0.0, v = 0.0;
    for (int q = -half; q <= half; q++) {
      const int j = fx.ns + q;
      const long idx = (long)k * M + (long)j * n1 + i;
      const cplx x = jb.psi[b][idx] * expi(-0.5 * jb.dt * jb.pot[idx]);
      d += fx.coef[q + half] * x;
      if (q == 0) v = x;
    }
    val[s] = v;
    der[s] = d;
  }
#pragma omp parallel for schedule(static)
  for (int e = 0; e < fx.ne; e++) {
    const cplx ph = expi(fx.E[e] * t) * jb.dt;
    cplx *a = fx.A.data() + (long)e * S, *bb = fx.B.data() + (long)e * S;
    for (long s = 0; s < S; s++) { a[s] += ph * val[s]; bb[s] += ph * der[s]; }
  }
}

static void flux_finish(Job &jb, Flux &fx, double *crp) {
  const int n1 = jb.n1;
  const long S = (long)jb.nO * jb.nt * n1;
  for (int e = 0; e < fx.ne; e++) {
    double im = 0.0;
    for (long s = 0; s < S; s++) {
      const int k = (s / n1) % jb.nt;
      im += jb.w[k] * (std::conj(fx.A[(long)e * S + s]) * fx.B[(long)e * S + s]).imag();
    }
    im *= jb.dr1;
    crp[e] = fx.eta[e] > 0 ? im / (2 * M_PI * jb.m2 * fx.eta[e]) : 0.0;
  }
}

// Central first-derivative weights on n = 2 half + 1 equispaced points (the
// stencil is exact for polynomials of degree <= 2 half):
//   c_m = (-1)^(m+1) (half!)^2 / (m (half-m)! (half+m)!) / h,  c_-m = -c_m, c_0 = 0
// evaluated as a product of ratios.  (Solving the Vandermonde system directly
// loses about 1e-7 at 21 points.)
static void fornberg(int n, double h, double *c) {
  const int half = n / 2;
  c[half] = 0.0;
  for (int m = 1; m <= half; m++) {
    double r = 1.0;
    for (int q = 0; q < m; q++) r *= (double)(half - q) / (double)(half + q + 1);
    const double cm = ((m % 2) ? 1.0 : -1.0) * r / m / h;
    c[half + m] = cm;
    c[half - m] = -cm;
  }
}

extern "C" void oracle_first_derivative(int n, double h, double *c) { fornberg(n, h, c); }

// oracle_evolve: same contract as qd_evolve (include/qd.h) for time.total_steps
// steps.  Returns 0 on success, nonzero with a message in err otherwise.
extern "C" int oracle_evolve(qd_inputs *in, int n_threads, char *err) {
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  err[0] = 0;
  Job jb;
  if (setup_job(in, jb, err)) return 1;
  const int nO = jb.nO;
  const long G = (long)jb.M * jb.nt;
  int &steps = *in->time.steps;
  const double dt = jb.dt;
  std::vector<double> ke(nO), re(nO), ce(nO), nrm(nO), pe(nO), c1(nO);
  Flux fx;
  const bool do_crp = in->crp.calculate_CRP && in->crp.n_energies > 0;
  if (do_crp) {
    fx.ns = in->crp.n_dividing_surface; fx.ng = in->crp.n_gradient_points; fx.ne = in->crp.n_energies;
    if (fx.ng < 3 || fx.ng % 2 == 0 || fx.ns - fx.ng / 2 < 0 || fx.ns + fx.ng / 2 >= jb.n2) {
      snprintf(err, 256, "bad flux surface"); return 1;
    }
    fx.coef.resize(fx.ng);
    fornberg(fx.ng, jb.dr2, fx.coef.data());
    fx.E.assign(in->crp.energies, in->crp.energies + fx.ne);
    fx.eta.assign(in->crp.eta_sq, in->crp.eta_sq + fx.ne);
    const long S = (long)nO * jb.nt * jb.n1;
    fx.A.assign(S * fx.ne, 0.0); fx.B.assign(S * fx.ne, 0.0);
  }
  const bool copy_enabled = (in->options.wave_to_matlab || in->options.wave_callback) &&
                            in->options.steps_to_copy_psi_from_device_to_host > 0;
  for (int s = 0; s < in->time.total_steps; s++) {
    if (jb.dump) dump_dvr(jb);
    if (steps == 0) phase_dvr(jb, -dt / 2);
    phase_dvr(jb, dt);
    legendre_forward(jb);
    rotational(jb, dt / 4);
    coriolis(jb, dt / 2, nullptr);
    rotational(jb, dt / 4);
    fft2d(jb, -1);
    kinetic(jb, dt);
    kinetic_energy(jb, ke.data());
    fft2d(jb, +1);
    scale_fbr(jb);
    rotational(jb, dt / 4);
    coriolis(jb, dt / 2, ce.data());
    rotational(jb, dt / 4);
    rotational_energy(jb, re.data());
    legendre_backward(jb);
    dvr_observables(jb, nrm.data(), pe.data());
    if (in->observables) {
      double *o = in->observables + (long)s * nO * QD_OBS_FIELDS;
      for (int b = 0; b < nO; b++) {
        o[b * QD_OBS_FIELDS + QD_OBS_NORM] = nrm[b];
        o[b * QD_OBS_FIELDS + QD_OBS_POTENTIAL] = pe[b];
        o[b * QD_OBS_FIELDS + QD_OBS_KINETIC] = ke[b];
        o[b * QD_OBS_FIELDS + QD_OBS_ROTATIONAL] = re[b];
        o[b * QD_OBS_FIELDS + QD_OBS_CORIOLIS] = ce[b];
      }
    }
    if (in->autocorrelation) {
      const cplx a = autocorrelation(jb);
      in->autocorrelation[2 * s] = a.real();
      in->autocorrelation[2 * s + 1] = a.imag();
    }
    if (do_crp) flux_accumulate(jb, fx, (steps + 1) * dt);
    double total = 0.0;
    for (int b = 0; b < nO; b++) total += nrm[b];
    if (total > 1.0 + 1e-10) {
      snprintf(err, 256, "total norm %.17g exceeds 1 at step %d", total, steps);
      return 3;
    }
    if (copy_enabled && steps % in->options.steps_to_copy_psi_from_device_to_host == 0) {
      cplx *dst = reinterpret_cast<cplx *>(in->omega_states.wave_packets);
      for (int b = 0; b < nO; b++) std::copy(jb.psi[b].begin(), jb.psi[b].end(), dst + b * G);
      if (in->options.wave_callback) in->options.wave_callback(in->options.callback_user, steps);
    }
    steps++;
  }
  if (in->final_copy) {
    cplx *dst = reinterpret_cast<cplx *>(in->omega_states.wave_packets);
    for (int b = 0; b < nO; b++) std::copy(jb.psi[b].begin(), jb.psi[b].end(), dst + b * G);
  }
  if (do_crp && in->crp.CRP) flux_finish(jb, fx, in->crp.CRP);
  return 0;
}
