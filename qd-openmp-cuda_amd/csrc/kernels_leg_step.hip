// kernels_leg_step.hip - the backward Legendre transform of one step fused
// with the forward Legendre transform of the next, on the matrix cores.
//
// Between the two transforms a step does only pointwise work on the DVR
// state: the potential energy of the stored state (omegawavepacket.cu:101-132),
// then, opening the next step, the absorber and exp(-i dt V)
// (cudaOpenMP2.cu:342-346; evolutionCUDAaux.cu:5-11, :243-252).  So one wave
// takes a strip of 16 points of one omega block from the FBR coefficients
// after step n to the FBR coefficients inside step n + 1:
//   psi  (16 x nt)  = FBR (16 x N_l) . P^T            (omegawavepacket.cu:221-240)
//   FBR' (16 x N_l) = (d1 d2 exp(-i dt V) o psi) . wP  (omegawavepacket.cu:200-219)
// and psi never goes to HBM and back: 32 bytes per DVR point and block less
// traffic per step, and one launch less.  In the paired-row MFMA layout the
// hand-off is register to register: item it of column tile t of the backward
// accumulators of lane (lr, lk) holds psi(point lr, theta 16 t + lk + 4 it),
// and the forward product's operand of k-step 4 t + it in the same lane is
// psi(point lr, theta 4 (4 t + it) + lk), the same value.
// Each product, the phase multiply and the potential-energy sum are the
// arithmetic of k_leg_bwd_paired and k_leg_fwd_paired operation for operation,
// so the FP64 state is the same to the bit as with the separate passes; only
// the partial sums of the potential energy are grouped by this kernel's
// workgroups.  The DVR values of the flux stencil rows are stored for the
// surface kernel.  Jobs whose state is in DVR (a new job, or after a separate
// backward pass) take the forward product alone.
#include "gemm_wave.h"

namespace qd {

// DVR rows stored by the fused kernel: the r rows [j0, j1) of block b (the
// flux stencil), all theta
struct StepRows {
  QD_G double2 *psi;   // block b, [nt][M]
  int n1, nt, j0, j1;
};

// B (a [kp][np] row-major table) into LDS rows of stride Ws: ROWS rows and
// W columns, zero past kp and np, so the fused products need no k-step
// guard (operand rows past K load as zero too).  Up to 8 loads per thread in
// flight before the LDS writes.
template <int TB, int ROWS, int W, typename Src>
__device__ __forceinline__ void stage_padded(double *Bs, Src Bg, int kp, int np, int Ws) {
  constexpr int CH = 8, total = ROWS * W;
  for (int e0 = 0; e0 < total; e0 += CH * TB) {
    double v[CH];
#pragma unroll
    for (int q = 0; q < CH; q++) {
      const int e = e0 + q * TB + (int)threadIdx.x;
      const int k = e / W, c = e - k * W;
      v[q] = (e < total && k < kp && c < np) ? Bg[k * np + c] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < CH; q++) {
      const int e = e0 + q * TB + (int)threadIdx.x;
      if (e < total) {
        const int k = e / W;
        Bs[k * Ws + (e - k * W)] = v[q];
      }
    }
  }
}

// One wave's strips of the fused step, with one wave per SIMD (the
// accumulators of both products and a four-deep operand ring take about 450
// registers).  A strip is a fixed sequence of tiles: the NT2 k-chunks of the
// backward product (FBR operand; bwd_kp / 16 = NT2 chunks), then the NT1
// k-chunks of the forward product (phase operand with the w V values of the
// backward tile it consumes; fwd_kp / 16 = NT1), padded with idle tiles to a
// multiple of lcm(D, 2).  The sequence is unrolled, so every register index
// is a compile-time constant (a register array indexed at run time goes to
// scratch memory), and fenced tile by tile (hoisted across the unrolled strip,
// the tiles' B fragments would all be live at once).  Within a fence, tile
// tau's MFMAs issue together with the preparation of tile tau + 1: its B
// fragments read from LDS (the tables are padded to whole chunks, so no
// k-step guard), and its operand taken from the ring, either raw FBR values
// or backward outputs times the phase, with the potential energy (in the
// order of gemm_wave_bwd_paired) and the flux stencil rows of that tile.  The
// VALU and LDS work thus runs in the shadow of the MFMAs; operands and B
// fragments are double-buffered.  Global loads run D - 1 tiles ahead through
// a ring of D slots, across strip boundaries.  Loads, clamping and masking
// as in gemm_wave_bwd_paired / gemm_wave_fwd_paired.  The output overwrites
// the FBR operand in place: a strip's stores follow all of its operand loads,
// and no other wave reads its points.
// FIN: the backward product alone, finishing a step sequence; the forward
// tiles only sum the potential energy and store their backward outputs (the
// DVR state, every row), so the partial sums are those of the fused form.
template <int NT1, int NT2, int D, bool FIN, typename Stage>
__device__ __forceinline__ void step_wave_paired(const GemmWave &g1, const GemmWave &g2, const StepRows &sr,
                                                 long strip0, int nstrips, int sstride, double &ob, Stage &&stage) {
  constexpr int kCh = 4;
  constexpr int T0 = NT2 + NT1;
  constexpr int LCM = D % 2 == 0 ? D : 2 * D;
  constexpr int T = (T0 + LCM - 1) / LCM * LCM;
  constexpr int NB = NT1 > NT2 ? NT1 : NT2;
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const long npts = g1.rows / 2;
  int ns = 0;
  for (int s = 0; s < nstrips; s++)
    if ((strip0 + (long)s * sstride) * 16 < npts) ns = s + 1;
  if (ns <= 0) {
    stage();
    return;
  }
  struct Slot {
    v2d x[kCh];
    double w[4];
  };
  struct Op {
    double ar[kCh], ai[kCh];
    double bv[kCh][NB];
  };
  Slot ring[D];
  Op op[2];
  const __amdgpu_buffer_rsrc_t rA = make_rsrc(g1.A, (uint32_t)((unsigned long)g1.K * (unsigned long)g1.lda * 8ul));
  const __amdgpu_buffer_rsrc_t rT =
      make_rsrc((const void *)g2.vph, (uint32_t)((unsigned long)g2.K * (unsigned long)g2.lda * 8ul));
  const uint32_t kstride = (uint32_t)g1.lda * 32u;
  auto load = [&](int tau, long st, Slot &r) {
    const long p = st * 16 + lr;
    const long pc = p < npts ? p : npts - 1;
    if (tau < NT2) {
      const uint32_t v0 = ((uint32_t)(4 * tau * kCh + lk) * (uint32_t)g1.lda + 2u * (uint32_t)pc) * 8u;
#pragma unroll
      for (int q = 0; q < kCh; q++) r.x[q] = buffer_load_v2d(rA, v0 + q * kstride);
    } else if (tau < T0) {
      const int c2 = tau - NT2;
      if (!FIN) {
        const uint32_t v0 = ((uint32_t)(4 * c2 * kCh + lk) * (uint32_t)g1.lda + 2u * (uint32_t)pc) * 8u;
#pragma unroll
        for (int q = 0; q < kCh; q++) r.x[q] = buffer_load_v2d(rT, v0 + q * kstride);
      }
#pragma unroll
      for (int it = 0; it < 4; it++) r.w[it] = g1.wpot[(long)acc_col<false>(c2, lk, it) * g1.M + pc];
    }
  };
  PairAcc<NT1> a1;
  PairAcc<NT2> a2;
  // operand and B fragments of tile tau of strip st (backward tiles read a1
  // only after the last backward MFMA of the strip, by data dependence)
  auto prep = [&](int tau, long st, bool surf, const Slot &r, Op &o) {
    const long p = st * 16 + lr;
    const bool pok = p < npts;
    if (tau < NT2) {
#pragma unroll
      for (int q = 0; q < kCh; q++) {
        o.ar[q] = r.x[q].x;
        o.ai[q] = r.x[q].y;
#pragma unroll
        for (int t = 0; t < NT1; t++) o.bv[q][t] = g1.Bs[(16 * tau + 4 * q + lk) * g1.Ws + lr + 16 * t];
      }
    } else if (tau < T0) {
      const int c2 = tau - NT2;
      const v4d pr = a1.r[c2], pi = a1.i[c2];
#pragma unroll
      for (int q = 0; q < kCh; q++) {
        ob += (pok ? pr[q] * pr[q] + pi[q] * pi[q] : 0.0) * r.w[q];   // w V = 0 for col >= nt
        const int col = acc_col<false>(c2, lk, q);
        if (surf && col < sr.nt) *(QD_G v2d *)(sr.psi + (long)col * g1.M + p) = v2d{pr[q], pi[q]};
        if (!FIN) {
          const v2d tv = r.x[q];
          o.ar[q] = fma(-pi[q], tv.y, pr[q] * tv.x);
          o.ai[q] = fma(pr[q], tv.y, pi[q] * tv.x);
#pragma unroll
          for (int t = 0; t < NT2; t++) o.bv[q][t] = g2.Bs[(16 * c2 + 4 * q + lk) * g2.Ws + lr + 16 * t];
        }
      }
    }
  };
  auto mma = [&](int tau, long st, const Op &o) {
    if (tau < NT2) {
      if (tau == 0) a1.reset();
#pragma unroll
      for (int q = 0; q < kCh; q++)
#pragma unroll
        for (int t = 0; t < NT1; t++) {
          a1.r[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(o.bv[q][t], o.ar[q], a1.r[t], 0, 0, 0);
          a1.i[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(o.bv[q][t], o.ai[q], a1.i[t], 0, 0, 0);
        }
    } else if (!FIN && tau < T0) {
      const int c2 = tau - NT2;
      if (c2 == 0) a2.reset();
#pragma unroll
      for (int q = 0; q < kCh; q++)
#pragma unroll
        for (int t = 0; t < NT2; t++) {
          a2.r[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(o.bv[q][t], o.ar[q], a2.r[t], 0, 0, 0);
          a2.i[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(o.bv[q][t], o.ai[q], a2.i[t], 0, 0, 0);
        }
      if (c2 == NT1 - 1) {   // streaming 16-byte stores, as the forward pass
        const long p = st * 16 + lr;
        const bool pok = p < npts;
#pragma unroll
        for (int t = 0; t < NT2; t++)
#pragma unroll
          for (int it = 0; it < 4; it++) {
            const int col = acc_col<false>(t, lk, it);
            if (pok && col < g2.N)
              __builtin_nontemporal_store(v2d{a2.re(t, it), a2.im(t, it)}, (v2d *)(g2.C + (long)col * g2.ldc + 2 * p));
          }
      }
    }
  };
  auto on_surface = [&](long st) {   // this lane's point lies on a flux stencil row (r row p / n1)
    const int pj = (int)(st * 16 + lr) / sr.n1;
    return st * 16 + lr < npts && pj >= sr.j0 && pj < sr.j1;
  };
#pragma unroll
  for (int tau = 0; tau < D - 1; tau++) load(tau, strip0, ring[tau]);
  stage();
  bool surf = on_surface(strip0);
  prep(0, strip0, surf, ring[0], op[0]);
  for (int s = 0; s < ns; s++) {
    const long st = strip0 + (long)s * sstride;
    const long stn = s + 1 < ns ? st + sstride : st;   // after the last strip: its own tiles again (unused)
    const bool surf_n = on_surface(stn);
#pragma unroll
    for (int tau = 0; tau < T; tau++) {
      const int tn = tau + D - 1;
      if (tn < T) load(tn, st, ring[tn % D]);
      else load(tn - T, stn, ring[tn % D]);
      mma(tau, st, op[tau % 2]);
      if (tau + 1 < T) prep(tau + 1, st, surf, ring[(tau + 1) % D], op[(tau + 1) % 2]);
      else prep(0, stn, surf_n, ring[0], op[0]);
      __builtin_amdgcn_sched_barrier(0);
    }
    surf = surf_n;
  }
}

// Workgroup (block b, group g) of a job: 4 waves, one per SIMD, the strips
// g S 4 + wave + 4 s of block b.  LDS: wP of the block, then (fused) P^T.
// FIN: the separate backward transform that finishes a fused step sequence
// (jobs without a pending backward transform do nothing).
template <int NT1, int NT2, bool FIN>
__global__ __launch_bounds__(4 * 64) void k_leg_step_paired(StepArgs a) {
  constexpr int WAVES = 4, TB = WAVES * 64;
  constexpr int Ws1 = lds_stride(NT1), Ws2 = lds_stride(NT2);
  extern __shared__ double lds_d[];
  __shared__ double red[WAVES];
  const int lu = xcd_logical_unit(blockIdx.x, gridDim.x);
  const int job = locate_job(a.unit_base, a.n_jobs, lu);
  const JobDev &jb = a.jobs[job];
  const int u = lu - a.unit_base[job];
  const int b = u % jb.nO, g = u / jb.nO;
  const int om = jb.omin + b, nl = jb.lmax - om + 1;
  const long M = jb.M;
  const long G = M * jb.nt;
  const long slab0 = ((long)b * jb.L + om) * M;
  const bool fused = jb.counters[3] != 0;   // the job's backward transform is pending
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // wP [16 NT1][Ws2] (theta rows, zero past fwd_kp), then P^T [16 NT2][Ws1]
  // (l rows, zero past bwd_kp)
  double *B2 = lds_d;
  double *B1 = lds_d + 16 * NT1 * Ws2;
  auto stage = [&]() {
    stage_padded<TB, 16 * NT1, NT2 * 16>(B2, jb.wP + (long)b * jb.fwd_kp * jb.fwd_np, jb.fwd_kp, jb.fwd_np, Ws2);
    if (fused)
      stage_padded<TB, 16 * NT2, NT1 * 16>(B1, jb.Pt + (long)b * jb.bwd_kp * jb.bwd_np, jb.bwd_kp, jb.bwd_np, Ws1);
    __syncthreads();
  };
  GemmWave g2{};
  g2.A = (const double *)(jb.psi + b * G);
  g2.C = (double *)(jb.fbr + slab0);
  g2.rows = 2 * M;
  g2.lda = g2.ldc = 2 * M;
  g2.K = jb.nt;
  g2.N = nl;
  g2.ksteps = jb.fwd_kp / 4;
  g2.Bs = B2;
  g2.Ws = Ws2;
  g2.vph = (const QD_G double *)jb.vphase;
  g2.wpot = jb.wpot;
  g2.M = M;
  const int S = jb.step_strips;
  const long strip0 = (long)g * S * WAVES + wave;
  if (!fused) {
    if (!FIN) gemm_wave_fwd_paired<NT2, 4, 3, false>(g2, strip0, S, WAVES, stage);
  } else {
    GemmWave g1{};
    g1.A = (const double *)(jb.fbr + slab0);
    g1.rows = 2 * M;
    g1.lda = g1.ldc = 2 * M;
    g1.K = nl;
    g1.N = jb.nt;
    g1.ksteps = jb.bwd_kp / 4;
    g1.Bs = B1;
    g1.Ws = Ws1;
    g1.wpot = jb.wpot;
    g1.M = M;
    StepRows sr;
    sr.psi = jb.psi + b * G;
    sr.n1 = jb.n1;
    sr.nt = jb.nt;
    // rows of the DVR state to store: the flux stencil, or (FIN) all
    sr.j0 = FIN ? 0 : jb.flux_on ? jb.flux_ns - jb.flux_ng / 2 : 0;
    sr.j1 = FIN ? jb.n2 : jb.flux_on ? sr.j0 + jb.flux_ng : 0;
    double ob = 0.0;
    step_wave_paired<NT1, NT2, 4, FIN>(g1, g2, sr, strip0, S, WAVES, ob, stage);
    ob = wave_sum64(ob);
    if (lane == 0) red[wave] = ob;
    __syncthreads();
    if (threadIdx.x == 0) jb.part_obs[(long)b * jb.step_groups + g] = (red[0] + red[1]) + (red[2] + red[3]);
  }
  // apply the counter advance of the reduction before this kernel (at the
  // end, so the store cannot hold back the descriptor loads above)
  if (u == 0 && threadIdx.x == 0) advance_counters(jb);
}

// QD_LEG_STEP=0 keeps the separate Legendre passes (tests, comparisons)
bool leg_step_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("QD_LEG_STEP");
    return !(e && std::atoi(e) == 0);
  }();
  return on && leg_paired();
}

bool leg_step_supported(int theta_tiles, int l_tiles, bool mixed) {
  return leg_step_enabled() && !mixed && theta_tiles >= 1 && theta_tiles <= 4 && l_tiles >= 1 &&
         l_tiles <= theta_tiles;
}

size_t leg_step_lds_bytes(int theta_tiles, int l_tiles) {
  return ((size_t)16 * l_tiles * lds_stride(theta_tiles) + (size_t)16 * theta_tiles * lds_stride(l_tiles)) *
         sizeof(double);
}

#define QD_STEP_CASES(X) X(1, 1) X(2, 1) X(2, 2) X(3, 1) X(3, 2) X(3, 3) X(4, 1) X(4, 2) X(4, 3) X(4, 4)

void launch_leg_step(const StepArgs &a, int total_units, int theta_tiles, int l_tiles, size_t lds, int finish,
                     hipStream_t s) {
  switch (theta_tiles * 8 + l_tiles) {
#define QD_STEP_LAUNCH(T, L)                                                                            \
  case T * 8 + L:                                                                                       \
    if (finish) hipLaunchKernelGGL((k_leg_step_paired<T, L, true>), dim3(total_units), dim3(256), lds, s, a); \
    else hipLaunchKernelGGL((k_leg_step_paired<T, L, false>), dim3(total_units), dim3(256), lds, s, a);       \
    break;
    QD_STEP_CASES(QD_STEP_LAUNCH)
#undef QD_STEP_LAUNCH
    default: break;   // excluded by leg_step_supported
  }
}

void init_attrs_leg_step() {
  const int lim = 160 * 1024;
#define QD_STEP_ATTR(T, L)                                                                                  \
  (void)hipFuncSetAttribute((const void *)k_leg_step_paired<T, L, false>,                                   \
                            hipFuncAttributeMaxDynamicSharedMemorySize, lim);                               \
  (void)hipFuncSetAttribute((const void *)k_leg_step_paired<T, L, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  QD_STEP_CASES(QD_STEP_ATTR)
#undef QD_STEP_ATTR
}

}  // namespace qd
