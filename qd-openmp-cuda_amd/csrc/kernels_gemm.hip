// kernels_gemm.hip - Legendre DVR <-> FBR transforms on the matrix cores,
// with the DVR pointwise work fused into their operand load and epilogue.
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
// Forward prologue: each A element (re or im of psi at (R_i, r_j, theta_k))
// is multiplied by d1[i] d2[j] exp(-i dt V) before it enters the MFMA; the
// re/im partner sits in the neighbouring lane.  (The extra exp(+i dt/2 V) of
// step 0 is applied once by k_prephase before the first step.)
// This is the absorber and potential phase of cudaOpenMP2.cu:342-346
// (evolutionCUDAaux.cu:5-11, :243-252), which the reference runs as separate
// passes over psi.
// Backward epilogue: the per-workgroup partial sums of (w_k V) |psi|^2, the
// potential energy of omegawavepacket.cu:101-132, from the accumulators (the
// norm and autocorrelation come from the FBR coefficients in the FFT pass).
//
// The mixed form splits both operands into FP32 hi/lo parts and accumulates
// hi*hi and hi*lo + lo*hi with v_mfma_f32_16x16x4f32 in FP32, flushing the FP32
// accumulators into FP64 every 16 k-terms.
#include <cstdlib>

#include "gemm_wave.h"

namespace qd {

// WAVES waves per workgroup share one LDS copy of B: 8 where B is too large
// for two 4-wave workgroups per CU.
template <int NT, int KC, bool FWD, bool MIXED, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void k_leg(StepArgs a) {
  constexpr int TB = WAVES * 64;
  extern __shared__ double lds_d[];
  __shared__ double red[WAVES];
  const int lu = xcd_logical_unit(blockIdx.x, gridDim.x);
  const int job = locate_job(a.unit_base, a.n_jobs, lu);
  const JobDev &jb = a.jobs[job];
  const int u = lu - a.unit_base[job];
  const int groups = FWD ? jb.gemm_fwd_groups : jb.gemm_bwd_groups;
  // consecutive workgroups take the same rows of consecutive blocks, so the
  // per-row tables (phase, potential) are reused from L2
  const int b = u % jb.nO, g = u / jb.nO;
  const int om = jb.omin + b, nl = jb.lmax - om + 1;
  const long M = jb.M;
  const long G = M * jb.nt;
  const long slab0 = ((long)b * jb.L + om) * M;
  const int kp = FWD ? jb.fwd_kp : jb.bwd_kp;
  const int np = FWD ? jb.fwd_np : jb.bwd_np;
  constexpr int Ws = lds_stride(NT);
  constexpr int W = NT * 16;
  // the wave index is uniform over the wave: readfirstlane tells the
  // compiler, so strip and tile indices (and the operand row bases) are scalar
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  double *Bs = lds_d;
  float *Bh = (float *)lds_d;
  v4u *Bx = (v4u *)(Bh + kp * Ws);
  const int xrows = 4 * ((kp + 15) / 16);
  auto stage = [&]() {
    if (!MIXED) {
      stage_b<TB, W>(Bs, (FWD ? jb.wP : jb.Pt) + (long)b * kp * np, kp, np, Ws);
    } else {
      stage_b<TB, W>(Bh, (FWD ? jb.wP_hi : jb.Pt_hi) + (long)b * kp * np, kp, np, Ws);
      stage_b<TB, W>(Bx, (const QD_G v4u *)(FWD ? jb.wP_x : jb.Pt_x) + (long)b * xrows * np, xrows, np, W);
    }
    __syncthreads();
  };
  GemmWave gw;
  gw.A = FWD ? (const double *)(jb.psi + b * G) : (const double *)(jb.fbr + slab0);
  gw.C = FWD ? (double *)(jb.fbr + slab0) : (double *)(jb.psi + b * G);
  gw.rows = 2 * M;
  gw.lda = gw.ldc = 2 * M;
  gw.K = FWD ? jb.nt : nl;
  gw.N = FWD ? nl : jb.nt;
  gw.ksteps = kp / 4;
  gw.Bs = Bs; gw.Bh = Bh; gw.Bl = nullptr; gw.Bx = Bx; gw.Ws = Ws;
  gw.vph = (const QD_G double *)jb.vphase;
  gw.wpot = jb.wpot; gw.M = M;
  double ob = 0.0;
  const int S = FWD ? jb.gemm_strips : jb.gemm_bwd_strips;
  gemm_wave<NT, KC, (KC == 8 && (NT <= 4 || (NT <= 6 && !MIXED))) ? 3 : 2, MIXED, FWD, !FWD>(
      gw, (long)g * S * WAVES + wave, S, WAVES, ob, stage);
  if (!FWD) {
    ob = wave_sum64(ob);
    if (lane == 0) red[wave] = ob;
    __syncthreads();
    if (threadIdx.x == 0) {
      double sum = (red[0] + red[1]) + (red[2] + red[3]);
      if constexpr (WAVES == 8) sum += (red[4] + red[5]) + (red[6] + red[7]);
      jb.part_obs[(long)b * groups + g] = sum;
    }
  }
  // apply the counter advance of the last reduction (at the end, so the
  // store cannot hold back the descriptor loads above); after the forward
  // pass of the same step the backward pass finds none
  if (u == 0 && threadIdx.x == 0) advance_counters(jb);
}

// Forward Legendre pass on paired rows (strips of 16 complex points).
template <int NT, int WAVES, bool MIXED>
__global__ __launch_bounds__(WAVES * 64) void k_leg_fwd_paired(StepArgs a) {
  constexpr int TB = WAVES * 64;
  extern __shared__ double lds_d[];
  const int lu = xcd_logical_unit(blockIdx.x, gridDim.x);
  const int job = locate_job(a.unit_base, a.n_jobs, lu);
  const JobDev &jb = a.jobs[job];
  const int u = lu - a.unit_base[job];
  const int b = u % jb.nO, g = u / jb.nO;
  const int om = jb.omin + b, nl = jb.lmax - om + 1;
  const long M = jb.M;
  const long G = M * jb.nt;
  const long slab0 = ((long)b * jb.L + om) * M;
  const int kp = jb.fwd_kp, np = jb.fwd_np;
  constexpr int Ws = lds_stride(NT);
  constexpr int W = NT * 16;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double *Bs = lds_d;
  float *Bh = (float *)lds_d;
  v4u *Bx = (v4u *)(Bh + kp * Ws);
  const int xrows = 4 * ((kp + 15) / 16);
  auto stage = [&]() {
    if (!MIXED) {
      stage_b<TB, W>(Bs, jb.wP + (long)b * kp * np, kp, np, Ws);
    } else {
      stage_b<TB, W>(Bh, jb.wP_hi + (long)b * kp * np, kp, np, Ws);
      stage_b<TB, W>(Bx, (const QD_G v4u *)jb.wP_x + (long)b * xrows * np, xrows, np, W);
    }
    __syncthreads();
  };
  GemmWave gw;
  gw.A = (const double *)(jb.psi + b * G);
  gw.C = (double *)(jb.fbr + slab0);
  gw.rows = 2 * M;
  gw.lda = gw.ldc = 2 * M;
  gw.K = jb.nt;
  gw.N = nl;
  gw.ksteps = kp / 4;
  gw.Bs = Bs; gw.Bh = Bh; gw.Bl = nullptr; gw.Bx = Bx; gw.Ws = Ws;
  gw.vph = (const QD_G double *)jb.vphase;
  gw.wpot = jb.wpot; gw.M = M;
  const int S = jb.gemm_strips;
  // mixed: two tiles in flight (three spill beside the FP32 partial sums)
  gemm_wave_fwd_paired<NT, 4, MIXED ? 2 : 3, MIXED>(gw, (long)g * S * WAVES + wave, S, WAVES, stage);
  // first kernel of the step: apply the previous step's counter advance
  if (u == 0 && threadIdx.x == 0) advance_counters(jb);
}

// Backward Legendre pass on paired rows, with the potential-energy partials.
template <int NT, int WAVES, bool MIXED>
__global__ __launch_bounds__(WAVES * 64) void k_leg_bwd_paired(StepArgs a) {
  constexpr int TB = WAVES * 64;
  extern __shared__ double lds_d[];
  __shared__ double red[WAVES];
  const int lu = xcd_logical_unit(blockIdx.x, gridDim.x);
  const int job = locate_job(a.unit_base, a.n_jobs, lu);
  const JobDev &jb = a.jobs[job];
  const int u = lu - a.unit_base[job];
  const int groups = jb.gemm_bwd_groups;
  const int b = u % jb.nO, g = u / jb.nO;
  const int om = jb.omin + b, nl = jb.lmax - om + 1;
  const long M = jb.M;
  const long G = M * jb.nt;
  const long slab0 = ((long)b * jb.L + om) * M;
  const int kp = jb.bwd_kp, np = jb.bwd_np;
  constexpr int Ws = lds_stride(NT);
  constexpr int W = NT * 16;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  double *Bs = lds_d;
  float *Bh = (float *)lds_d;
  v4u *Bx = (v4u *)(Bh + kp * Ws);
  const int xrows = 4 * ((kp + 15) / 16);
  auto stage = [&]() {
    if (!MIXED) {
      stage_b<TB, W>(Bs, jb.Pt + (long)b * kp * np, kp, np, Ws);
    } else {
      stage_b<TB, W>(Bh, jb.Pt_hi + (long)b * kp * np, kp, np, Ws);
      stage_b<TB, W>(Bx, (const QD_G v4u *)jb.Pt_x + (long)b * xrows * np, xrows, np, W);
    }
    __syncthreads();
  };
  GemmWave gw;
  gw.A = (const double *)(jb.fbr + slab0);
  gw.C = (double *)(jb.psi + b * G);
  gw.rows = 2 * M;
  gw.lda = gw.ldc = 2 * M;
  gw.K = nl;
  gw.N = jb.nt;
  gw.ksteps = kp / 4;
  gw.Bs = Bs; gw.Bh = Bh; gw.Bl = nullptr; gw.Bx = Bx; gw.Ws = Ws;
  gw.vph = (const QD_G double *)jb.vphase;
  gw.wpot = jb.wpot; gw.M = M;
  double ob = 0.0;
  const int S = jb.gemm_bwd_strips;
  // mixed: two tiles in flight (three spill beside the epilogue's potentials)
  gemm_wave_bwd_paired<NT, 4, MIXED ? 2 : 3, MIXED>(gw, (long)g * S * WAVES + wave, S, WAVES, ob, stage);
  ob = wave_sum64(ob);
  if (lane == 0) red[wave] = ob;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = (red[0] + red[1]) + (red[2] + red[3]);
    if constexpr (WAVES == 8) sum += (red[4] + red[5]) + (red[6] + red[7]);
    jb.part_obs[(long)b * groups + g] = sum;
  }
  // the counter advance of the last reduction (a separate backward pass that
  // finishes a fused step sequence follows a reduction directly)
  if (u == 0 && threadIdx.x == 0) advance_counters(jb);
}

// LDS of the Legendre kernels: B as FP64, or (mixed) B hi as FP32 plus the
// bfloat16 compensation table of 16 bytes per chunk of 16 k, lane group and column
size_t leg_lds_bytes(int kp, int np, int mixed) {
  const int nt = np / 16;
  if (!mixed) return (size_t)kp * lds_stride(nt) * sizeof(double);
  return (size_t)kp * lds_stride(nt) * sizeof(float) + (size_t)4 * ((kp + 15) / 16) * np * 16;
}

template <int NT>
static void launch_leg_bwd_paired_nt(const StepArgs &a, int total_units, int kp_max, int mixed, int waves,
                                     hipStream_t s) {
  const size_t lds = leg_lds_bytes(kp_max, NT * 16, mixed);
  if (mixed) {
    if (waves == 8) hipLaunchKernelGGL((k_leg_bwd_paired<NT, 8, true>), dim3(total_units), dim3(512), lds, s, a);
    else hipLaunchKernelGGL((k_leg_bwd_paired<NT, 4, true>), dim3(total_units), dim3(256), lds, s, a);
  } else {
    if (waves == 8) hipLaunchKernelGGL((k_leg_bwd_paired<NT, 8, false>), dim3(total_units), dim3(512), lds, s, a);
    else hipLaunchKernelGGL((k_leg_bwd_paired<NT, 4, false>), dim3(total_units), dim3(256), lds, s, a);
  }
}

// FP64 Legendre passes on paired rows (up to 4 column tiles: wider ones spill,
// or with strips of 16 points fill fewer CUs than the one-row kernels);
// QD_LEG_PAIRED=0 selects the one-row-per-lane kernels.  Read once per
// process, so every job agrees.  On the J=20 workload the paired kernels take
// 63 / 57 us against 68 / 66 us.
bool leg_paired() {
  static const bool v = [] {
    const char *e = std::getenv("QD_LEG_PAIRED");
    return !(e && std::atoi(e) == 0);
  }();
  return v;
}

template <int NT>
static void launch_leg_fwd_paired_nt(const StepArgs &a, int total_units, int kp_max, int mixed, int waves,
                                     hipStream_t s) {
  const size_t lds = leg_lds_bytes(kp_max, NT * 16, mixed);
  if (mixed) {
    if (waves == 8) hipLaunchKernelGGL((k_leg_fwd_paired<NT, 8, true>), dim3(total_units), dim3(512), lds, s, a);
    else hipLaunchKernelGGL((k_leg_fwd_paired<NT, 4, true>), dim3(total_units), dim3(256), lds, s, a);
  } else {
    if (waves == 8) hipLaunchKernelGGL((k_leg_fwd_paired<NT, 8, false>), dim3(total_units), dim3(512), lds, s, a);
    else hipLaunchKernelGGL((k_leg_fwd_paired<NT, 4, false>), dim3(total_units), dim3(256), lds, s, a);
  }
}

template <int NT, int KC, bool FWD, int WAVES>
static void launch_leg_k(const StepArgs &a, int total_units, int kp_max, int mixed, hipStream_t s) {
  const size_t lds = leg_lds_bytes(kp_max, NT * 16, mixed);
  if (mixed) hipLaunchKernelGGL((k_leg<NT, KC, FWD, true, WAVES>), dim3(total_units), dim3(WAVES * 64), lds, s, a);
  else hipLaunchKernelGGL((k_leg<NT, KC, FWD, false, WAVES>), dim3(total_units), dim3(WAVES * 64), lds, s, a);
}

// QD_LEG_KC=8|12|16 overrides the k-chunk choice (tuning sweeps)
static int leg_chunk_override() {
  static const int v = [] {
    const char *e = std::getenv("QD_LEG_KC");
    const int k = e ? std::atoi(e) : 0;
    return (k == 8 || k == 12 || k == 16) ? k : 0;
  }();
  return v;
}

template <int NT, bool FWD>
static void launch_leg_nt(const StepArgs &a, int total_units, int kp_max, int mixed, int waves, hipStream_t s) {
  if (waves == 8) {   // large B: KC = 8 only
    launch_leg_k<NT, 8, FWD, 8>(a, total_units, kp_max, mixed, s);
    return;
  }
  const int kc = leg_chunk_override() ? leg_chunk_override() : leg_chunk(kp_max / 4, FWD, NT, mixed != 0);
  switch (NT <= 8 ? kc : 8) {
    case 16: launch_leg_k<NT, NT <= 8 ? 16 : 8, FWD, 4>(a, total_units, kp_max, mixed, s); break;
    case 12: launch_leg_k<NT, NT <= 8 ? 12 : 8, FWD, 4>(a, total_units, kp_max, mixed, s); break;
    default: launch_leg_k<NT, 8, FWD, 4>(a, total_units, kp_max, mixed, s); break;
  }
}

template <bool FWD>
static void launch_leg(const StepArgs &a, int total_units, int nt_tiles, int kp_max, int mixed, int waves,
                       hipStream_t s) {
  switch (nt_tiles) {
    case 1: launch_leg_nt<1, FWD>(a, total_units, kp_max, mixed, waves, s); break;
    case 2: launch_leg_nt<2, FWD>(a, total_units, kp_max, mixed, waves, s); break;
    case 3: launch_leg_nt<3, FWD>(a, total_units, kp_max, mixed, waves, s); break;
    case 4: launch_leg_nt<4, FWD>(a, total_units, kp_max, mixed, waves, s); break;
    case 5: launch_leg_nt<5, FWD>(a, total_units, kp_max, mixed, waves, s); break;
    case 6: launch_leg_nt<6, FWD>(a, total_units, kp_max, mixed, waves, s); break;
    case 7: launch_leg_nt<7, FWD>(a, total_units, kp_max, mixed, waves, s); break;
    case 8: launch_leg_nt<8, FWD>(a, total_units, kp_max, mixed, waves, s); break;
    case 10: launch_leg_nt<10, FWD>(a, total_units, kp_max, mixed, waves, s); break;
    case 12: launch_leg_nt<12, FWD>(a, total_units, kp_max, mixed, waves, s); break;
    case 15: launch_leg_nt<15, FWD>(a, total_units, kp_max, mixed, waves, s); break;
    default: break;   // rejected by leg_tiles_supported at validation
  }
}

bool leg_tiles_supported(int nt_tiles) {
  switch (nt_tiles) {
    case 1: case 2: case 3: case 4: case 5: case 6: case 7: case 8: case 10: case 12: case 15: return true;
    default: return false;
  }
}

// Paired-row kernels: up to 4 column tiles in both passes; the FP64 forward
// pass also at 5 and 6 tiles for jobs of two or more blocks (9 blocks of
// 256 x 192 x 96: 373 -> 340 us).  One block of 256 x 192 x 96 has too few
// strips of 16 points to fill the CUs (59 against 55 us), the backward
// pass spills beyond 4 tiles and so does the mixed forward pass.
bool leg_fwd_paired(int nt_tiles, bool mixed, int n_blocks) {
  return leg_paired() && (nt_tiles <= 4 || (!mixed && nt_tiles <= 6 && n_blocks >= 2));
}
bool leg_bwd_paired(int nt_tiles) { return leg_paired() && nt_tiles <= 4; }

void launch_leg_fwd(const StepArgs &a, int total_units, int nt_tiles, int kp_max, int mixed, int waves, int paired,
                    hipStream_t s) {
  if (paired) {
    switch (nt_tiles) {
      case 1: launch_leg_fwd_paired_nt<1>(a, total_units, kp_max, mixed, waves, s); return;
      case 2: launch_leg_fwd_paired_nt<2>(a, total_units, kp_max, mixed, waves, s); return;
      case 3: launch_leg_fwd_paired_nt<3>(a, total_units, kp_max, mixed, waves, s); return;
      case 4: launch_leg_fwd_paired_nt<4>(a, total_units, kp_max, mixed, waves, s); return;
      case 5: launch_leg_fwd_paired_nt<5>(a, total_units, kp_max, mixed, waves, s); return;
      case 6: launch_leg_fwd_paired_nt<6>(a, total_units, kp_max, mixed, waves, s); return;
      default: break;
    }
  }
  launch_leg<true>(a, total_units, nt_tiles, kp_max, mixed, waves, s);
}
void launch_leg_bwd(const StepArgs &a, int total_units, int nt_tiles, int kp_max, int mixed, int waves, int paired,
                    hipStream_t s) {
  if (paired) {   // 5 and 6 tiles spill in 8-wave workgroups
    switch (nt_tiles) {
      case 1: launch_leg_bwd_paired_nt<1>(a, total_units, kp_max, mixed, waves, s); return;
      case 2: launch_leg_bwd_paired_nt<2>(a, total_units, kp_max, mixed, waves, s); return;
      case 3: launch_leg_bwd_paired_nt<3>(a, total_units, kp_max, mixed, waves, s); return;
      case 4: launch_leg_bwd_paired_nt<4>(a, total_units, kp_max, mixed, waves, s); return;
      default: break;
    }
  }
  launch_leg<false>(a, total_units, nt_tiles, kp_max, mixed, waves, s);
}

// probe: C (rows x n, ldc) = A (rows x k, lda) . B ([kp][np] row-major), one strip per wave
template <int NT>
__global__ __launch_bounds__(kBlock) void k_gemm_probe(const double *A, int lda, const double *Bg, int kp, int np,
                                                       double *C, int ldc, int rows, int K, int N) {
  extern __shared__ double Bs[];
  constexpr int Ws = lds_stride(NT), W = NT * 16;
  const int r0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
  GemmWave gw{};
  gw.A = A; gw.lda = lda; gw.K = K; gw.ksteps = kp / 4;
  gw.C = C; gw.ldc = ldc; gw.N = N; gw.rows = rows;
  gw.Bs = Bs; gw.Ws = Ws;
  double ob = 0.0;
  gemm_wave<NT, 8, 2, false, false, false>(gw, r0 / 16, 1, 1, ob, [&]() {
    stage_b<kBlock, W>(Bs, Bg, kp, np, Ws);
    __syncthreads();
  });
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

template <int NT, int KC, int WAVES = 4>
static void set_leg_attrs_k(int lim) {
  (void)hipFuncSetAttribute((const void *)k_leg<NT, KC, true, false, WAVES>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_leg<NT, KC, false, false, WAVES>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_leg<NT, KC, true, true, WAVES>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_leg<NT, KC, false, true, WAVES>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
}

template <int NT>
static void set_leg_attrs(int lim) {
  set_leg_attrs_k<NT, 8>(lim);
  set_leg_attrs_k<NT, 8, 8>(lim);
  if (NT <= 8) {
    set_leg_attrs_k<NT, NT <= 8 ? 12 : 8>(lim);
    set_leg_attrs_k<NT, NT <= 8 ? 16 : 8>(lim);
  }
}

template <int NT, bool MIXED>
static void set_leg_paired_attrs_p(int lim) {
  (void)hipFuncSetAttribute((const void *)k_leg_fwd_paired<NT, 4, MIXED>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_leg_fwd_paired<NT, 8, MIXED>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_leg_bwd_paired<NT, 4, MIXED>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  (void)hipFuncSetAttribute((const void *)k_leg_bwd_paired<NT, 8, MIXED>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
}
template <int NT>
static void set_leg_paired_attrs(int lim) {
  set_leg_paired_attrs_p<NT, false>(lim);
  set_leg_paired_attrs_p<NT, true>(lim);
}

void init_attrs_gemm() {
  const int lim = 160 * 1024;
  set_leg_attrs<1>(lim); set_leg_attrs<2>(lim); set_leg_attrs<3>(lim); set_leg_attrs<4>(lim);
  set_leg_attrs<5>(lim); set_leg_attrs<6>(lim); set_leg_attrs<7>(lim); set_leg_attrs<8>(lim);
  set_leg_attrs<10>(lim); set_leg_attrs<12>(lim); set_leg_attrs<15>(lim);
  (void)hipFuncSetAttribute((const void *)k_gemm_probe<8>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  set_leg_paired_attrs<1>(lim); set_leg_paired_attrs<2>(lim); set_leg_paired_attrs<3>(lim);
  set_leg_paired_attrs<4>(lim); set_leg_paired_attrs<5>(lim); set_leg_paired_attrs<6>(lim);
}

}  // namespace qd
