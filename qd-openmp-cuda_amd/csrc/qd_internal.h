// qd_internal.h - device-side job descriptor and kernel entry points shared by
// the HIP translation units of libqdamd.
//
// One JobDev describes one partial-wave job (all of its omega blocks).  Every
// step kernel takes an array of JobDev plus a per-kernel prefix table that
// maps blockIdx.x to (job, local work unit), so a single launch can advance
// one job (qd_run) or every admitted job of the server at once.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Device pointers of the job descriptor are global-memory pointers.  Marking
// them so in the device compile lets the compiler emit global_load/store
// (instead of flat accesses, whose waits also drain the LDS counter); the host
// compile sees plain pointers with the same layout.
#if defined(__HIP_DEVICE_COMPILE__)
#define QD_G __attribute__((address_space(1)))
#define QD_C __attribute__((address_space(4)))
#else
#define QD_G
#define QD_C
#endif

namespace qd {

constexpr int kMaxOmega = 24;     // omega blocks per job (K_max <= 23)
constexpr int kFftMaxN = 512;     // longest FFT pencil
constexpr int kMaxTheta = 512;    // theta points per job
constexpr int kBlock = 256;       // threads per workgroup for the step kernels
constexpr int kRedFields = 7;     // fields of the final reduction: ke, ce, re, norm, acf re, acf im, pe

// kernels of a step that take work units (index into JobDev::units and the
// engine's prefix tables); the absorber / potential phase, the rotational-
// Coriolis half steps and the DVR observables are fused into these
enum KernelKind {
  KK_LEG_FWD = 0,
  KK_FFT_R,     // forward R pass
  KK_FFT_RI,    // inverse R pass (its own rows per workgroup)
  KK_FFT_r,
  KK_LEG_BWD,
  KK_REDUCE,
  KK_FLUX,
  KK_LEG_STEP,  // backward Legendre transform of one step fused with the forward one of the next
  KK_COUNT
};

struct JobDev {
  // shape
  int n1, n2, nt, M, L, lmax, nO, omin, kmax;
  int J;
  int mixed;                 // 1: mixed-precision Legendre GEMMs
  // work decomposition
  int gemm_strips, gemm_bwd_strips;   // strips per wave in the forward / backward Legendre GEMM
                                      // (16 rows, or 16 complex points in the paired-row passes)
  int gemm_fwd_waves, gemm_bwd_waves;   // waves per GEMM workgroup (4, or 8 to share a large B in LDS)
  int leg_fwd_paired, leg_bwd_paired;   // 1: the Legendre pass runs the paired-row kernels
  int gemm_fwd_groups;       // workgroups per omega block (forward GEMM)
  int gemm_bwd_groups;       // workgroups per omega block (backward GEMM)
  int fwd_kp, fwd_np;        // padded K (mult of 4) and N (mult of 16), forward
  int bwd_kp, bwd_np;        // padded K and N, backward
  int fft_rows, fft_rows_inv; // pencils per workgroup along R (forward, inverse R pass)
  int fft_cols;              // pencils per workgroup along r
  // two-pass form of the FFT work (1 when used): an r pass that applies the
  // first rotational . Coriolis . rotational half step and the r kinetic
  // factor (mix_cols columns of every block per workgroup), then an R pass
  // with the R kinetic factor and the second half step; otherwise three
  // passes (forward R with the first half step, r with both kinetic factors,
  // inverse R with the second half step)
  int fft2, mix_cols;
  // fused Legendre step (KK_LEG_STEP): 1 when the job's steps run it; its
  // waves per workgroup, strips (16 points) per wave and workgroups per block
  int leg_step, step_waves, step_strips, step_groups;
  // parity-split form of the fused step (theta grid and tables symmetric
  // under x -> -x): 1 when used; tiles of 16 for the half theta grid and the
  // even / odd l columns
  int leg_split, sp_th, sp_te, sp_to;
  // FFT passes store their output with streaming stores: set when one
  // wavepacket array (16 M nt nO bytes) is larger than the 256 MB Infinity
  // Cache; smaller ones are left there for the next pass to read
  int stream_out;
  int units[KK_COUNT];       // work units (workgroups) per kernel for this job
  // physics
  double dt, dr1, dr2;
  // state
  QD_G double2 *psi;              // [nO][nt][M] DVR wavepackets
  QD_G double2 *fbr;              // [nO][L][M]  FBR wavepackets (slabs l < omega unused)
  QD_G double2 *c0;               // [nO][L][M] FBR coefficients of the initial state, or null
  // tables
  const QD_G double2 *vphase;     // [nt][M]  d1[i] d2[j] exp(-i dt V)
  const QD_G double *pot;         // [nt][M]
  const QD_G double *w;           // [nt] Gauss-Legendre weights (flux surface integral)
  const QD_G double *wpot;        // [bwd_np][M] w_k V(theta_k, m), zero rows k >= nt
  const QD_G double *wP;          // [nO][fwd_kp][fwd_np] row-major (c fastest): w_k P_l(x_k)
  const QD_G double *Pt;          // [nO][bwd_kp][bwd_np] row-major: P_l(x_k), k = l - omega
  const QD_G float *wP_hi, *Pt_hi;// mixed form: FP32 hi parts of wP and Pt
  // mixed form: per block [ceil(kp / 16) * 4][np] 16-byte entries,
  // entry (4 c + g, n) = bfloat16 (B_lo(k_q, n), q < 4; B_hi(k_q, n), q < 4)
  // with k_q = 16 c + g + 4 q and B_hi = fp32(B), B_lo = fp32(B - B_hi)
  const QD_G uint4 *wP_x, *Pt_x;
  // parity-split tables, padded to whole tiles with zeros (k < nt / 2,
  // m: l = omega + 2 m (E) or omega + 2 m + 1 (O)):
  //   sp_PE [nO][16 sp_te][16 sp_th] = P_l(x_k),  sp_PO [nO][16 sp_to][16 sp_th]
  //   sp_wPE [nO][16 sp_th][16 sp_te] = w_k P_l(x_k),  sp_wPO [nO][16 sp_th][16 sp_to]
  const QD_G double *sp_PE, *sp_PO, *sp_wPE, *sp_wPO;
  // mixed form of the split tables: FP32 hi parts (same layout) and bfloat16
  // compensation tables (bf16_compensation_table layout, k = the table's rows).
  // The forward tables' rows are permuted within each 16-row chunk: row
  // 16 c + 4 q + g holds node 16 c + 4 g + q (the FP32 MFMA's output order)
  const QD_G float *sp_hPE, *sp_hPO, *sp_hwPE, *sp_hwPO;
  const QD_G uint4 *sp_xPE, *sp_xPO, *sp_xwPE, *sp_xwPO;
  const QD_G double *corV;        // [L][nO][nO] eigenvectors of C_l: corV[(l nO + a) nO + b] = v_a(omega_b)
  const QD_G double2 *corPhi;     // [L][n1][nO] exp(-i dt/2 l(l+1) I1_i) exp(-i dt/2 e_a I1_i)
  const QD_G double2 *rot2;       // [L][n2] exp(-i dt/2 l(l+1) / (2 m2 r_j^2))
  const QD_G double2 *kin1;       // [n1] exp(-i dt k_i^2 / (2 m1))
  const QD_G double2 *kin2;       // [n2] exp(-i dt k_j^2 / (2 m2))
  const QD_G double *kinE1;       // [n1] k_i^2/(2 m1)
  const QD_G double *kinE2;       // [n2]
  const QD_G double *I1;          // [n1] 1/(2 m1 R_i^2)
  const QD_G double *I2;          // [n2] 1/(2 m2 r_j^2)
  const QD_G double2 *cmat;       // [L][nO] (C_l(omega,omega), C_l(omega,omega+1)) exact entries
  const QD_G double2 *twR;        // [n1] exp(-2 pi i e / n1)
  const QD_G double2 *twr;        // [n2] exp(-2 pi i e / n2)
  // reduction scratch (per-workgroup partial sums)
  QD_G double *part_cor;          // [nO][5][units FFT_R] Coriolis, rotational energy, norm, acf re/im
  QD_G double *part_ke;           // [units FFT_r]       kinetic energy of one (block, slab, column group);
                                  // two-pass form: [nO][units FFT_RI] per R workgroup and block
  QD_G double *part_ke2;          // two-pass form: [nO][units FFT_r] r-kinetic sums of the r pass
  QD_G double *part_obs;          // [nO][gemm_bwd_groups or step_groups] w V |psi|^2 (potential energy)
  QD_G double *obs_ring;          // [ring][nO][5] per-step observables
  QD_G double2 *acf_ring;         // [ring][nO] autocorrelation contribution of each block
  int ring;                  // ring capacity (records)
  QD_G int *counters;             // device: [0] step counter (the reference's time.steps),
                             //         [1] records written, [2] advance pending: set by
                             //         the step's reduction, applied by advance_counters
                             //         at the end of the next Legendre kernel;
                             //         [3] backward transform pending (leg_step jobs): set
                             //         by the inverse R pass, cleared by the ungated
                             //         reduction that follows a separate backward pass
  // flux at the dividing surface r2 = r_s (index flux_ns): per step the
  // surface kernel writes the values and the ng-point r-derivative of the
  // state with V(dt/2) restored into window slot (step % flux_T); every
  // flux_T steps the window is folded into the energy accumulators
  int flux_on, flux_ns, flux_ng, flux_ne, flux_T;
  int flux_S;                     // surface points per step: nO * nt * n1
  const QD_G double *flux_coef;   // [ng] first-derivative weights
  const QD_G double2 *flux_rph;   // [ng][nt][n1] exp(-i dt/2 V) on the stencil rows
  QD_G double2 *flux_W;           // [T][2][S] window: values, derivatives
};

// Arguments of the windowed time-to-energy transform of one job:
//   A[e][s] += sum_q ph[e][q] W[(s0 + q) % T][0][s],  B likewise with W[.][1]
//   ph[e][q] = exp(i E_e t_q) dt, t_q = (s0 + q + 1) dt
struct FluxWindow {
  const double2 *W;
  double2 *A, *B;                 // [ne][S]
  double2 *ph;                    // [ne][T]
  const double *E;                // [ne]
  double dt;
  int S, ne, T, s0, n;
};

// Per-launch arguments common to all step kernels.
struct StepArgs {
  const QD_G JobDev *jobs;
  const QD_G int *unit_base;      // [n_jobs + 1] prefix of units for this kernel kind
  int n_jobs;
  // flux surface and reduction kernels: 1 in the fused step sequence, where
  // they finish the previous step of each job whose backward transform was
  // pending (counters[3]) and skip the others; 0 after a separate backward pass
  int gate;
  // first FFT pass of the fused step sequence (the forward R pass, or the
  // r pass with the first half step in the two-pass form): its first red_units
  // workgroups run the (gated) reduction of the previous step, with the
  // reduction's prefix table red_base (0: none)
  const QD_G int *red_base;
  int red_units;
  // ... and its next flux_units workgroups the (gated) flux surface values of
  // that step, with the surface kind's prefix table flux_base (0: none)
  const QD_G int *flux_base;
  int flux_units;
};

// Locate the job owning workgroup `u` (binary search over the prefix table).
__device__ __forceinline__ int locate_job(const QD_G int *base, int n_jobs, int u) {
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= u) lo = mid; else hi = mid - 1;
  }
  return lo;
}

#if defined(__HIPCC__)   // device helpers below need the amdgcn builtins (not in the g++ host objects)
// Lane permutation of a double within rows of 16 lanes (DPP control CTRL):
// a VALU move, no LDS traffic.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)(v & 0xffffffffLL), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(v >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double readlane_f64(double x, int lane) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum of x over the 64 lanes of the wave, the same (uniform) value in every
// lane; all lanes must be active.  Within each row of 16 lanes: add the quad
// neighbours (lane ^ 1, lane ^ 2), then the row rotated by 4 and by 8; then
// the four row sums, read from lanes 0, 16, 32, 48, in a fixed order.  DPP
// moves and lane reads only (no LDS permutes), and a fixed order of every
// addition, so the result is reproducible bit for bit.
__device__ __forceinline__ double wave_sum64(double x) {
  x += dpp_f64<0xB1>(x);    // quad_perm [1, 0, 3, 2]
  x += dpp_f64<0x4E>(x);    // quad_perm [2, 3, 0, 1]
  x += dpp_f64<0x124>(x);   // row_ror 4
  x += dpp_f64<0x128>(x);   // row_ror 8
  return (readlane_f64(x, 0) + readlane_f64(x, 16)) + (readlane_f64(x, 32) + readlane_f64(x, 48));
}
#endif

// Applies the step-counter advance flagged by the last reduction; called,
// after its own work, by one thread of the job's first workgroup of each
// Legendre kernel (forward, fused or backward; none reads a counter).  A
// reduction is always followed by such a kernel before the next flux surface
// kernel or reduction, and kernel boundaries order the advance between them;
// a second call finds nothing pending.
__device__ __forceinline__ void advance_counters(const JobDev &jb) {
  if (jb.counters[2]) {
    jb.counters[0] += 1;
    jb.counters[1] += 1;
    jb.counters[2] = 0;
  }
}

// Host-side launchers (each enqueues one kernel on `stream`).
// step-0 phase exp(+i dt/2 V) on every block of one job (before its first step)
void launch_prephase(double2 *psi, const double *pot, long G, int nO, double dt, hipStream_t stream);
void launch_wpot(double *wpot, const double *pot, const double *w, int nt, long M, int rows, hipStream_t stream);
void launch_vphase(double2 *vph, const double *pot, const double *d1, const double *d2, int n1, int n2, long G,
                   double dt, hipStream_t stream);
// column tile counts (N / 16) with a Legendre GEMM instantiation
bool leg_tiles_supported(int nt_tiles);
// Legendre passes on paired rows (default; QD_LEG_PAIRED=0 turns them off)
bool leg_paired();
// whether a job's forward / backward Legendre pass runs the paired-row kernels
bool leg_fwd_paired(int nt_tiles, bool mixed, int n_blocks);
bool leg_bwd_paired(int nt_tiles);
// LDS bytes of a Legendre kernel for B of kp x np (np a multiple of 16)
size_t leg_lds_bytes(int kp, int np, int mixed);
void launch_leg_fwd(const StepArgs &a, int total_units, int nt_tiles, int kp_max, int mixed, int waves, int paired,
                    hipStream_t stream);
// Fused Legendre step (FP64, paired rows): for each job whose backward
// transform is pending (counters[3]), the backward product of its last step,
// the potential-energy partials, the absorber and potential phase and the
// forward product of its next step, strip by strip, the DVR state held in
// registers; for the others the forward product alone.  theta_tiles = bwd_np
// / 16 and l_tiles = fwd_np / 16 (both supported by leg_step_supported).
bool leg_step_enabled();
bool leg_step_supported(int theta_tiles, int l_tiles, bool mixed);
// Parity-split form of the fused kernel (kernels_leg_split.hip): supported
// tile counts (half theta grid, even and odd l), LDS bytes, launch (finish as
// launch_leg_step); mixed: the FP32 hi / bfloat16 compensation form.
bool leg_split_supported(int th, int te, int to, bool mixed);
size_t leg_split_lds_bytes(int th, int te, int to, int mixed);
void launch_leg_split(const StepArgs &a, int total_units, int th, int te, int to, int mixed, size_t lds, int finish,
                      hipStream_t stream);
// LDS bytes of the fused kernel: both Legendre tables, padded to whole
// chunks of 16 rows
size_t leg_step_lds_bytes(int theta_tiles, int l_tiles);
// finish: the backward transform alone, for the jobs with one pending (the
// separate backward pass that ends a fused step sequence, with the same
// partial sums of the potential energy)
void launch_leg_step(const StepArgs &a, int total_units, int theta_tiles, int l_tiles, size_t lds, int finish,
                     hipStream_t stream);
void launch_leg_bwd(const StepArgs &a, int total_units, int nt_tiles, int kp_max, int mixed, int waves, int paired,
                    hipStream_t stream);
// FFT along R fused with rotational . Coriolis . rotational (before the forward
// transform, after the inverse one; the inverse pass also forms the energies),
// and the r pass with the kinetic phase.  _rt: any length factoring into
// 2, 3, 5 (dynamic LDS); _ct: compile-time lengths (fft_ct_supported), with
// fft_ct_rows(n1, blocks, inverse) rows / fft_ct_cols(n2) columns per
// workgroup (the rows depend on the number of omega blocks and on the pass).
void launch_fft_R_rt(const StepArgs &a, int total_units, int inverse, int acf, int nb_max, size_t lds, hipStream_t stream);
void launch_fft_r_rt(const StepArgs &a, int total_units, size_t lds, hipStream_t stream);
bool fft_ct_supported(int n);
int fft_ct_rows(int n1, int n_blocks, bool inverse);
int fft_ct_cols(int n2);
void launch_fft_R_ct(const StepArgs &a, int total_units, int inverse, int acf, int n1, int nb_max, hipStream_t stream);
void launch_fft_r_ct(const StepArgs &a, int total_units, int n2, hipStream_t stream);
// Two-pass form (compile-time lengths, up to 5 omega blocks): columns per
// workgroup of the r pass with the first half step (0: no such pass for these
// sizes, or turned off with QD_FFT_TWO_PASS=0), that r pass, and the R pass
// with the R kinetic factor and the second half step (fft_ct_rows(n1,
// blocks, true) rows per workgroup).
int fft_mix_cols(int n2, int n_blocks);
void launch_fft_r_mix(const StepArgs &a, int total_units, int n2, int nb_max, hipStream_t stream);
void launch_fft_R_kin(const StepArgs &a, int total_units, int acf, int n1, int nb_max, hipStream_t stream);
// standalone batched 1-D FFT of `count` contiguous pencils (test probe)
void launch_fft_probe(double2 *data, const double2 *tw, int n, int count, int inverse, hipStream_t stream);
void launch_reduce(const StepArgs &a, int total_units, hipStream_t stream);
// surface values of every flux job (one step, part of the step sequence)
void launch_flux(const StepArgs &a, int total_units, hipStream_t stream);
// fold the window of one job into its energy accumulators
void launch_flux_window(const FluxWindow &w, hipStream_t stream);
// P(E_e) = Im(sum_s w_k conj(A) B) dr1 / (2 pi mu_r eta_e) for every energy
void launch_flux_finish(const double2 *A, const double2 *B, const double *w, const double *eta, double *out,
                        int S, int n1, int nt, int ne, double scale, hipStream_t stream);

// plain copy kernel (traffic calibration)
void launch_copy(const double2 *src, double2 *dst, long n, hipStream_t stream);
// HBM ceiling probes: kind 1 wide copy with nontemporal stores, kind 2 read only
// (dst receives 2048 partial sums)
void launch_bandwidth(int kind, const double2 *src, double2 *dst, long n, hipStream_t stream);

// GEMM probe for tests: C (rows x n) = A (rows x k, lda) * B ([kp][np] row-major)
void launch_gemm_probe(const double *A, int lda, const double *B, int kp, int np, double *C, int ldc,
                       int rows, int k, int n, hipStream_t stream);

}  // namespace qd
