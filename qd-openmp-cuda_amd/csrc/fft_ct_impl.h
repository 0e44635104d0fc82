// fft_ct_impl.h - compile-time-length FFT passes for the common grid
// sizes (32, 48, 64, 96, 128, 192, 256 points along R or r).
//
// Same operators as kernels_fft.hip (see there for the reference
// counterparts).  With the length, the pencils per workgroup and the radix
// plan known at compile time, every index computation divides by constants
// and the stages unroll.  Each stage is in-place on one LDS tile: a thread
// reads the inputs of its butterflies into registers, the workgroup
// synchronises, and the outputs are written back, so a tile needs no
// ping-pong buffer.  In the R pass the tiles of all omega blocks present at
// slab l stay in registers (the Coriolis mix couples blocks point by point)
// and only one block at a time passes through LDS.
#pragma once
#include <type_traits>

#include "fft_common.h"
#include "step_reduce.h"
#include "flux_surface.h"

namespace qd {

// Radix plans.  PlanStd: 4s first, then 2, 3s, 5s.  PlanEnds: the first and
// the last stage share the radix R1 (8 when 64 divides n, else 4; n must be a
// multiple of R1^2) with the standard plan in between, so the last stage of
// a transform leaves each thread the points (idx + q n/R1) that the first
// stage of the next transform takes.
struct PlanStd {
  static constexpr int radix(int n, int ns) { return next_radix(n / ns); }
};
struct PlanEnds {
  static constexpr int first(int n) { return n % 64 == 0 ? 8 : 4; }
  static constexpr int radix(int n, int ns) {
    return (ns == 1 || ns * first(n) == n) ? first(n) : next_radix(n / (ns * first(n)));
  }
};
// PlanEnds with radix 4 at both ends (n a multiple of 16), for the r pass
// that holds every block's points in registers: 4 per butterfly and block;
// in between radix 8 where it divides what remains (one LDS round trip for
// 4 8 4 = 128), else the standard plan.
struct PlanEnds4 {
  static constexpr int first(int) { return 4; }
  static constexpr int radix(int n, int ns) {
    return (ns == 1 || ns * 4 == n) ? 4 : next_radix(n / (ns * 4));
  }
};

// R-pass LDS layout (SW).  16-byte LDS stores are served 8 lanes at a time
// over 128 bytes (32 banks), loads 16 lanes at a time over 256 bytes; a
// wave's access is conflict-free when every group spreads over distinct
// 16-byte slots (slot mod 8 for stores, mod 16 for loads), measured with
// SQ_LDS_BANK_CONFLICT on fixed lane patterns.
//  - Tile element i sits in slot swz(i): bits 0-1 XOR bits 3-4 and bit 2 XOR
//    bit 4.  The stores of the radix-4 stages, at lane stride 4 (stride-1
//    stage: 4 i + q) and in runs of 4 at stride 16 (stride-4 stage:
//    16 (i / 4) + i % 4 + 4 q), which unswizzled fall on 2 and 4 of the 8
//    slots, cover all 8; each aligned group of 8 elements keeps its slots, so
//    the unit-stride loads (every stage's inputs, the last stage, the point
//    map) stay conflict-free.
//  - Twiddles in stage order (tw_stage_offset): stage NS holds w^(q k N / (NS R))
//    at (q - 1) NS + k, so a stage reads them at unit stride in k instead of at
//    stride q N / (NS R).
template <bool SW> __device__ __forceinline__ int swz(int i) {
  return SW ? i ^ ((i >> 3) & 3) ^ (((i >> 4) & 1) << 2) : i;
}
// Offset of stage NS's twiddles in the stage-ordered table (stages of
// stride 1 have none); tw_stage_offset(n, n) is the table's length.
template <class PLAN> __host__ __device__ constexpr int tw_stage_offset(int n, int ns_stop) {
  int off = 0;
  for (int ns = 1; ns < ns_stop; ns *= PLAN::radix(n, ns))
    if (ns > 1) off += (PLAN::radix(n, ns) - 1) * ns;
  return off;
}
// Index into the natural twiddle table (w^j, j < N) of entry e of the
// stage-ordered one (stages unrolled at compile time).
template <class PLAN, int N, int NS = 1, int OFF = 0>
__device__ __forceinline__ int tw_stage_source(int e) {
  if constexpr (NS >= N) {
    return 0;
  } else {
    constexpr int R = PLAN::radix(N, NS);
    if constexpr (NS == 1) {
      return tw_stage_source<PLAN, N, NS * R, OFF>(e);
    } else {
      if (e < OFF + (R - 1) * NS) {
        const int d = e - OFF;
        return (d / NS + 1) * (d % NS) * (N / (NS * R));
      }
      return tw_stage_source<PLAN, N, NS * R, OFF + (R - 1) * NS>(e);
    }
  }
}

// Stockham stages NS.. (up to, not including, the stage of stride NSEND) of
// an in-place transform of NPEN pencils; pencil p element e is
// A[p * PS + e * ES] (slot swz<SW> of it).  PF maps consecutive threads to
// consecutive pencils.  Entry and exit are synchronised.
template <bool INV, int N, int NS, int NPEN, int PS, int ES, bool PF, class PLAN = PlanStd, int NSEND = N,
          bool SW = false>
__device__ __forceinline__ void stages_ip(double2 *A, const double2 *tw) {
  if constexpr (NS < NSEND) {
    constexpr int R = PLAN::radix(N, NS);
    constexpr int NB = N / R;
    constexpr int TOTAL = NPEN * NB;
    constexpr int MB = (TOTAL + kBlock - 1) / kBlock;
    constexpr int TWS = N / (NS * R);
    constexpr int TWO = tw_stage_offset<PLAN>(N, NS);
    double2 X[MB][R];
#pragma unroll
    for (int m = 0; m < MB; m++) {
      const int t = threadIdx.x + m * kBlock;
      if (TOTAL % kBlock == 0 || t < TOTAL) {
        const int p = PF ? t % NPEN : t / NB;
        const int idx = PF ? t / NPEN : t % NB;
        double2 v[R];
#pragma unroll
        for (int q = 0; q < R; q++) v[q] = A[swz<SW>(p * PS + (idx + q * NB) * ES)];
        if constexpr (NS > 1) {
          const int k = idx % NS;
#pragma unroll
          for (int q = 1; q < R; q++) {
            double2 w = SW ? tw[TWO + (q - 1) * NS + k] : tw[q * k * TWS];
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
        for (int q = 0; q < R; q++) A[swz<SW>(p * PS + (d + q * NS) * ES)] = X[m][q];
      }
    }
    __syncthreads();
    stages_ip<INV, N, NS * R, NPEN, PS, ES, PF, PLAN, NSEND, SW>(A, tw);
  }
}

// Orders a wave's LDS accesses across lanes (a wave's LDS instructions run
// in order; the fences keep the compiler from moving accesses across).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-local form of stages_ip for a row-major tile A[p * N + e] whose
// pencils belong to single waves: wave w owns pencils w + 4 j (j < PPW) and
// runs their butterflies on its lanes, 64 / NB pencils side by side when a
// stage has NB <= 32 butterflies per pencil, so the stages need no workgroup
// barrier, only the wave's own ordering.  Same plan (PlanStd), twiddles and
// slots as stages_ip.  Entry: the wave's pencils written and ordered
// (wave_lds_sync); exit likewise.
template <bool INV, int N, int NS, int PPW, bool SW, int NSEND = N>
__device__ __forceinline__ void stages_wave(double2 *A, const double2 *tw) {
  if constexpr (NS < NSEND) {
    constexpr int R = PlanStd::radix(N, NS);
    constexpr int NB = N / R;
    static_assert(NB <= 64, "one butterfly per lane and pencil");
    constexpr int SIDE = 64 / NB < PPW ? 64 / NB : PPW;   // pencils side by side on the lanes
    constexpr int PASSES = (PPW + SIDE - 1) / SIDE;
    constexpr int TWS = N / (NS * R);
    constexpr int TWO = tw_stage_offset<PlanStd>(N, NS);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int idx = lane % NB, js = lane / NB;   // butterfly and pencil slot of this lane
    if (js < SIDE) {
      double2 X[PASSES][R];
#pragma unroll
      for (int m = 0; m < PASSES; m++) {
        const int j = m * SIDE + js;
        const int p = wave + 4 * (j < PPW ? j : 0);
        double2 v[R];
#pragma unroll
        for (int q = 0; q < R; q++) v[q] = A[swz<SW>(p * N + idx + q * NB)];
        if constexpr (NS > 1) {
          const int k = idx % NS;
#pragma unroll
          for (int q = 1; q < R; q++) {
            double2 w = SW ? tw[TWO + (q - 1) * NS + k] : tw[q * k * TWS];
            if (INV) w.y = -w.y;
            v[q] = cmul(v[q], w);
          }
        }
        dft_small<INV, R>(v, X[m]);
      }
      // every lane's reads of the stage precede its writes in the wave's
      // instruction stream, so the stage needs no ordering in between
      const int d = (idx / NS) * NS * R + idx % NS;
#pragma unroll
      for (int m = 0; m < PASSES; m++) {
        const int j = m * SIDE + js;
        if (j < PPW) {
          const int p = wave + 4 * j;
#pragma unroll
          for (int q = 0; q < R; q++) A[swz<SW>(p * N + d + q * NS)] = X[m][q];
        }
      }
    }
    wave_lds_sync();
    stages_wave<INV, N, NS * R, PPW, SW, NSEND>(A, tw);
  }
}

// Stride of the last stage of PlanStd for length n.
__host__ __device__ constexpr int last_stage_ns(int n) {
  int ns = 1;
  while (ns * next_radix(n / ns) < n) ns *= next_radix(n / ns);
  return ns;
}

// Last Stockham stage (stride NS, radix R = N / NS, twiddles w^(q idx)) of the
// transforms of the NPEN pencils of a row-major tile A[p * N + e], into
// registers: butterfly t = tid + m kBlock (p = t / NS, idx = t % NS) reads
// and yields the points p N + idx + q NS, stored at y[m R + q] (threads past
// the last butterfly compute pencil 0 again; their values are not used).
// Its inputs are exactly its outputs' positions, so a thread that writes
// back only these positions needs no barrier before it.  Entry must be
// synchronised.
template <bool INV, int N, int NS, int NPEN, int XPT, bool SW = false>
__device__ __forceinline__ void last_stage_own(const double2 *A, const double2 *tw, double2 (&y)[XPT]) {
  constexpr int R = N / NS;
  constexpr int TWO = tw_stage_offset<PlanStd>(N, NS);
  constexpr int MBL = XPT / R;
  constexpr int TL = NPEN * NS;
#pragma unroll
  for (int m = 0; m < MBL; m++) {
    const int t0 = threadIdx.x + m * kBlock;
    const int t = (TL % kBlock == 0 || t0 < TL) ? t0 : 0;
    const int p = t / NS, idx = t % NS;
    double2 v[R], X[R];
#pragma unroll
    for (int q = 0; q < R; q++) v[q] = A[swz<SW>(p * N + idx + q * NS)];
    if constexpr (NS > 1) {
#pragma unroll
      for (int q = 1; q < R; q++) {
        double2 w = SW ? tw[TWO + (q - 1) * NS + idx] : tw[q * idx];
        if (INV) w.y = -w.y;
        v[q] = cmul(v[q], w);
      }
    }
    dft_small<INV, R>(v, X);
#pragma unroll
    for (int q = 0; q < R; q++) y[m * R + q] = X[q];
  }
}

// Rows per workgroup of the R passes: for n1 > 128, 4 (fewer rows lose more
// to the per-workgroup table staging than they gain in occupancy), or 2 / 1
// when more than 9 / 16 omega blocks (the inverse pass: more than 5) are held
// in registers across the transform (the inverse pass with 4 rows of 9
// blocks has one wave per SIMD); for shorter pencils 8, or 2 when more than
// 5 blocks are held.  Measured on 9 blocks of 192 x 128 x 64: forward 121 us
// with 4 rows and 160 us with 2, inverse 289 us with 4 and 174 us with 2.
__host__ __device__ constexpr int ct_rows(int n1, int nb, bool inv) {
  return n1 > 128 ? (nb <= (inv ? 5 : 9) ? 4 : (nb <= 16 ? 2 : 1)) : (nb <= 5 ? 8 : 2);
}
// Columns per workgroup of the r pass: 16 or 8, so the boundary stages of
// its radix plan have 128..256 butterflies for the 256 threads.
__host__ __device__ constexpr int ct_cols(int n2) { return (n2 <= 64 || n2 == 128) ? 16 : 8; }
// Radix of the end stages of the two-pass form's r pass: PlanEnds's 8 for
// one block on pencils of 128 or more points that 64 divides (on 192 x 256
// with 96 theta points: 44.5 -> 41.2 us; on 64 points, where half the
// threads would idle at the ends, 10.1 -> 11.2), else 4 (PlanEnds4): with
// more blocks all their points stay in registers through the Coriolis mix.
__host__ __device__ constexpr int mix_radix(int n2, int nb) {
  return nb <= 1 && n2 % 64 == 0 && n2 >= 128 ? 8 : 4;
}
// Columns per workgroup of that r pass (k_fft_r_mix): mix_radix points per
// butterfly and ceil(Q n2 / radix / kBlock) butterflies a thread, so Q is the
// larger of 16 and 8 with a tile of at most 2048 points and at most 20
// points a thread over the instantiated block count (1, 2, 3 or 5), one
// butterfly a thread with more than one block (two spill); 0 (the
// three-pass form runs) above 5 blocks or where neither fits.  Narrower
// tiles would read and write rows of less than 128 bytes.
__host__ __device__ constexpr int mix_cols(int n2, int nb) {
  if (nb > 5 || n2 % 16 != 0) return 0;
  const int b = nb <= 3 ? (nb < 1 ? 1 : nb) : 5;
  const int r1 = mix_radix(n2, b);
  for (int q = 16; q >= 8; q /= 2) {
    const int mb = (q * n2 / r1 + kBlock - 1) / kBlock;
    if (q * n2 <= 2048 && b * r1 * mb <= 20 && (b == 1 || mb == 1)) return q;
  }
  return 0;
}

// First stage of the Coriolis mix for the XPT points of a thread and NB
// blocks, in place: t_a(k) = phi(k, a) sum_b V[a][b] x_b(k).  Every entry
// V[a][b] (a, b < nO; the eigenvector table is zero past the blocks present
// at the slab) is read once and applied to all points, so one entry is live
// at a time where the per-point form keeps all NB^2 of them in registers.
template <int NB, int XPT, class VPtr, class Phi>
__device__ __forceinline__ void mix_first_stage(double2 (&x)[NB][XPT], VPtr V, int nO, Phi phi) {
  const int nOc = nO - 1;
  double2 t[NB][XPT];
#pragma unroll
  for (int al = 0; al < NB; al++) {
    double2 acc[XPT];
#pragma unroll
    for (int k = 0; k < XPT; k++) acc[k] = make_double2(0.0, 0.0);
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const double v = V[min(al, nOc) * nO + min(b, nOc)];
#pragma unroll
      for (int k = 0; k < XPT; k++) {
        acc[k].x = fma(v, x[b][k].x, acc[k].x);
        acc[k].y = fma(v, x[b][k].y, acc[k].y);
      }
    }
#pragma unroll
    for (int k = 0; k < XPT; k++) t[al][k] = cmul(phi(k, al), acc[k]);
  }
#pragma unroll
  for (int al = 0; al < NB; al++)
#pragma unroll
    for (int k = 0; k < XPT; k++) x[al][k] = t[al][k];
}

// ---------------------------------------------------------------------------
// R pass with the Coriolis mix (forward: mix then DFT; inverse: DFT, scale,
// energies, mix).  KIN (inverse only, the two-pass form): each block's
// forward DFT, the R kinetic factor kin1 with the R part of the kinetic
// energy, then the inverse DFT.
// ---------------------------------------------------------------------------
template <bool INV, int NB, int N1, bool ACF, bool KIN = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(KIN && NB == 5 && N1 == 192 ? 4 : 1))) void k_fft_R_ct(StepArgs a) {
  static_assert(INV || !KIN, "the R kinetic factor rides on the inverse pass");
  constexpr int NF = ACF ? 5 : 3;
  constexpr int P = ct_rows(N1, NB, INV);
  constexpr int PN = P * N1;
  // Points of this thread.  OWN: the outputs of its butterflies in the last
  // FFT stage (stride NSL, radix RL): point k = m RL + q of butterfly
  // t = tid + m kBlock (row p = t / NSL, idx = t % NSL) is e = p N1 + idx + q NSL.
  // Every per-point step (loads, Coriolis mix, energies, stores) uses this
  // map, so the last stage reads only this thread's points, runs in
  // registers and writes HBM (forward) or the mix operands (inverse)
  // directly.  Used where it gives each thread as many points as the plain
  // map (e = tid + k kBlock), i.e. where the last stage keeps every thread
  // busy; short pencils with few rows keep the plain map and a full LDS pass.
  constexpr int PPT = (PN + kBlock - 1) / kBlock;
  constexpr int NSL = last_stage_ns(N1);
  constexpr int RL = N1 / NSL;
  constexpr int TL = P * NSL, MBL = (TL + kBlock - 1) / kBlock;
  constexpr bool OWN = MBL * RL == PPT;
  // WV: with the last stage's stride 64 every point of a thread lies in a
  // pencil of its own wave (pencils wave + 4 m), so the FFT stages run wave
  // by wave (stages_wave) without workgroup barriers
  constexpr bool WV = OWN && NSL == 64 && P % 4 == 0;
  constexpr int XPT = PPT;
  auto ept = [](int k) {
    if constexpr (OWN) {
      const int t = threadIdx.x + (k / RL) * kBlock;
      return (t / NSL) * N1 + t % NSL + (k % RL) * NSL;
    } else {
      return (int)threadIdx.x + k * kBlock;
    }
  };
  auto eok = [](int k) {
    if constexpr (OWN) return TL % kBlock == 0 || (int)threadIdx.x + (k / RL) * kBlock < TL;
    else return PN % kBlock == 0 || (int)threadIdx.x + k * kBlock < PN;
  };
  // swizzled tile slots where the pencils are whole groups of 64 points
  constexpr bool SW = N1 % 64 == 0;
  constexpr int NTWS = SW ? tw_stage_offset<PlanStd>(N1, N1) : N1;
  __shared__ double2 A[PN];
  __shared__ double2 tw[NTWS];   // natural order, or stage order (SW)
  __shared__ double2 phs[N1 * NB];   // corPhi(l, i, a) of this slab, reused by the P rows
  __shared__ double red[4 * NF * NB];
  __shared__ double redk[KIN ? 4 * NB : 1];   // R kinetic sums per wave and block
  // inverse pass with more than 5 blocks: the slab's eigenvectors and C_l
  // entries in LDS (too many for SGPRs; from global memory they would be
  // reloaded after every store of the mix).  The forward mix has no stores
  // and reads them from global memory, where the compiler keeps them scalar.
  constexpr bool VLDS = INV && NB > 5;
  constexpr bool kMixBatched = INV && NB > 5 && NB <= 9 && N1 >= 256;
  __shared__ double Vs[VLDS ? NB * NB : 1];
  __shared__ double2 cms[VLDS ? NB : 1];
  if (!INV && reduce_carried(a, red)) return;   // leading workgroups: the previous step's reduction
  if (!INV && flux_carried(a)) return;          // then its flux surface values
  const int bid = blockIdx.x - (INV ? 0 : a.red_units + a.flux_units);
  const int job = locate_job(a.unit_base, a.n_jobs, bid);
  const JobDev &jb = a.jobs[job];
  const int stream = jb.stream_out;   // read before any store (stores may alias the descriptor)
  const int u = bid - a.unit_base[job];
  const int M = jb.M;
  const long LM = (long)jb.L * M;
  const int groups = jb.n2 / P;
  const int l = jb.omin + u / groups, g = u % groups;
  const int nb = min(jb.kmax, l) - jb.omin + 1;
  // uniform over the workgroup and read-only: scalar loads into SGPRs for up
  // to 5 blocks, an LDS copy (staged below; nO <= NB) for more in the inverse
  const auto V = [&]() {
    if constexpr (NB <= 5) return (const QD_C double *)(jb.corV + (long)l * jb.nO * jb.nO);
    else if constexpr (VLDS) return (const double *)Vs;
    else return (const QD_G double *)(jb.corV + (long)l * jb.nO * jb.nO);
  }();
  QD_G double2 *base = jb.fbr + (long)l * M + (long)g * PN;
  const QD_G double2 *c0 = ACF ? jb.c0 + (long)l * M + (long)g * PN : nullptr;
  // every global load of the slab tile in flight at once, issued before the
  // table staging below so that the round trips overlap; loads are
  // unconditional from in-bounds addresses (block nb-1, last point), and
  // values of blocks >= nb and points >= PN are never used.  The inverse
  // pass, which takes the blocks one at a time through its transforms before
  // the mix: the first PRE blocks only, and block b + PRE at the start of
  // block b, so the slab's loads spread over the workgroup's transform work
  // (192 x 128 x 64, 5 blocks: two-transform pass 77.5 -> 72.0 us, inverse
  // 52.0 -> 49.3; 9 blocks of 256 x 192 x 96: inverse 363.8 -> 348.8; three
  // blocks ahead: 80.9 with spills)
  constexpr int PRE = INV ? 2 : NB;
  double2 x[NB][XPT];
  auto load_block = [&](int b) {
#pragma unroll
    for (int k = 0; k < XPT; k++) x[b][k] = base[(b < nb ? b : nb - 1) * LM + (eok(k) ? ept(k) : 0)];
  };
#pragma unroll
  for (int b = 0; b < PRE && b < NB; b++) load_block(b);
  {
    // the twiddles and the slab's phases: every load first (clamped in-range
    // indices), then the LDS writes, so the staging costs one round trip
    // rather than one per loop trip
    constexpr int NTW = (NTWS + kBlock - 1) / kBlock;
    constexpr int NPH = (N1 * NB + kBlock - 1) / kBlock;
    const QD_G double2 *cp = jb.corPhi + (long)l * N1 * jb.nO;
    double2 tv[NTW], pv[NPH];
#pragma unroll
    for (int it = 0; it < NTW; it++) {
      const int e = min((int)threadIdx.x + it * kBlock, NTWS - 1);
      tv[it] = jb.twR[SW ? tw_stage_source<PlanStd, N1>(e) : e];
    }
#pragma unroll
    for (int it = 0; it < NPH; it++) {
      const int e = min((int)threadIdx.x + it * kBlock, N1 * NB - 1);
      pv[it] = cp[(e / NB) * jb.nO + min(e % NB, jb.nO - 1)];
    }
    if constexpr (VLDS) {
      for (int e = threadIdx.x; e < jb.nO * jb.nO; e += kBlock) Vs[e] = jb.corV[(long)l * jb.nO * jb.nO + e];
      for (int e = threadIdx.x; e < jb.nO; e += kBlock) cms[e] = jb.cmat[(long)l * jb.nO + e];
    }
#pragma unroll
    for (int it = 0; it < NTW; it++)
      if ((int)threadIdx.x + it * kBlock < NTWS) tw[threadIdx.x + it * kBlock] = tv[it];
#pragma unroll
    for (int it = 0; it < NPH; it++) {
      const int e = threadIdx.x + it * kBlock;
      if (e < N1 * NB) phs[e] = e % NB < jb.nO ? pv[it] : make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  FbrSums<NB, NF> sm;
  sm.zero();
  if (!INV) {
    {
#pragma unroll
      for (int k = 0; k < XPT; k++) {
        const int e = ept(k);
        if (eok(k)) {
          double2 xp[NB];
#pragma unroll
          for (int b = 0; b < NB; b++) xp[b] = x[b][k];
          coriolis_point<false, NB, NF>(jb, xp, nb, l, e % N1, g * P + e / N1, V, sm, nullptr, LM,
                                        (const double2 *)phs + (e % N1) * NB);
#pragma unroll
          for (int b = 0; b < NB; b++) x[b][k] = xp[b];
        }
      }
    }
#pragma unroll
    for (int b = 0; b < NB; b++) {
      if (b < nb) {
#pragma unroll
        for (int k = 0; k < XPT; k++)
          if (eok(k)) A[swz<SW>(ept(k))] = x[b][k];
        if constexpr (WV) wave_lds_sync();
        else __syncthreads();
        if constexpr (WV) {
          stages_wave<false, N1, 1, P / 4, SW, NSL>(A, tw);
          double2 y[XPT];
          last_stage_own<false, N1, NSL, P, XPT, SW>(A, tw, y);
#pragma unroll
          for (int k = 0; k < XPT; k++)
            if (eok(k)) store_out(base + b * LM + ept(k), y[k], stream);
          wave_lds_sync();   // the last stage's reads before the next block's writes
        } else if constexpr (OWN) {
          stages_ip<false, N1, 1, P, N1, 1, false, PlanStd, NSL, SW>(A, tw);
          double2 y[XPT];
          last_stage_own<false, N1, NSL, P, XPT, SW>(A, tw, y);
#pragma unroll
          for (int k = 0; k < XPT; k++)
            if (eok(k)) store_out(base + b * LM + ept(k), y[k], stream);
        } else {
          stages_ip<false, N1, 1, P, N1, 1, false, PlanStd, N1, SW>(A, tw);
#pragma unroll
          for (int k = 0; k < XPT; k++)
            if (eok(k)) store_out(base + b * LM + ept(k), A[swz<SW>(ept(k))], stream);
        }
      }
    }
  } else {
    const double s = 1.0 / (double)M;
    const int wave = threadIdx.x >> 6;
    constexpr int NV = NF * NB;
    // Blocks absent at this slab (b >= nb) hold x = 0, and the tables are
    // zero there (corV rows and columns, cmat entries and the last off-
    // diagonal, corPhi past nO), so the sums over blocks below run over all
    // NB with clamped, in-range table indices and no per-block conditions.
    const int nOc = jb.nO - 1;
    const auto cm = [&]() {
      if constexpr (NB <= 5) return (const QD_C double2 *)(jb.cmat + (long)l * jb.nO);
      else return (const double2 *)cms;
    }();
    // Coriolis energy of block b - 1 once block b is out of the transform:
    //   sum I1 (C(b,b) |x_b|^2 + 2 C(b,b+1) Re(conj(x_b) x_{b+1})),
    // wave-reduced at once so one sum is live at a time
    auto coriolis_energy = [&](int b) {
      const double2 c = cm[min(b, nOc)];
      double ce = 0.0;
#pragma unroll
      for (int k = 0; k < XPT; k++) {
        const int e = ept(k);
        if (eok(k)) {
          double en = c.x * cnorm(x[b][k]);
          if (b + 1 < NB) en += 2.0 * c.y * cdot_re(x[b][k], x[b + 1 < NB ? b + 1 : b][k]);
          ce += en * jb.I1[e % N1];
        }
      }
      wave_sum_to(ce, red + wave * NV + b * NF);
    };
    // blocks pass through LDS in turn
#pragma unroll
    for (int b = 0; b < NB; b++) {
      double ks = 0.0;   // KIN: R kinetic energy of this thread's points of block b
      if constexpr (INV) {
        if (b + PRE < NB) load_block(b + PRE);
      }
      if (b < nb) {
#pragma unroll
        for (int k = 0; k < XPT; k++)
          if (eok(k)) A[swz<SW>(ept(k))] = x[b][k];
        if constexpr (WV) wave_lds_sync();
        else __syncthreads();
        if constexpr (KIN) {
          // forward DFT; the thread's points (last-stage outputs under the
          // OWN map) take kin1 and go back to their own slots, so the
          // inverse DFT below needs only the usual entry ordering
          double2 y[XPT];
          if constexpr (WV) {
            stages_wave<false, N1, 1, P / 4, SW, NSL>(A, tw);
            last_stage_own<false, N1, NSL, P, XPT, SW>(A, tw, y);
          } else if constexpr (OWN) {
            stages_ip<false, N1, 1, P, N1, 1, false, PlanStd, NSL, SW>(A, tw);
            last_stage_own<false, N1, NSL, P, XPT, SW>(A, tw, y);
          } else {
            stages_ip<false, N1, 1, P, N1, 1, false, PlanStd, N1, SW>(A, tw);
#pragma unroll
            for (int k = 0; k < XPT; k++) y[k] = eok(k) ? A[swz<SW>(ept(k))] : make_double2(0.0, 0.0);
          }
#pragma unroll
          for (int k = 0; k < XPT; k++) {
            if (eok(k)) {
              const int i = ept(k) % N1;
              ks += cnorm(y[k]) * jb.kinE1[i];
              A[swz<SW>(ept(k))] = cmul(y[k], jb.kin1[i]);
            }
          }
          if constexpr (WV) wave_lds_sync();
          else __syncthreads();
        }
        if constexpr (WV) {
          stages_wave<true, N1, 1, P / 4, SW, NSL>(A, tw);
          last_stage_own<true, N1, NSL, P, XPT, SW>(A, tw, x[b]);
          wave_lds_sync();   // the last stage's reads before the next block's writes
        } else if constexpr (OWN) {
          stages_ip<true, N1, 1, P, N1, 1, false, PlanStd, NSL, SW>(A, tw);
          last_stage_own<true, N1, NSL, P, XPT, SW>(A, tw, x[b]);
        } else {
          stages_ip<true, N1, 1, P, N1, 1, false, PlanStd, N1, SW>(A, tw);
#pragma unroll
          for (int k = 0; k < XPT; k++)
            if (eok(k)) x[b][k] = A[swz<SW>(ept(k))];
        }
#pragma unroll
        for (int k = 0; k < XPT; k++) x[b][k] = eok(k) ? cscale(x[b][k], s) : make_double2(0.0, 0.0);
      } else {
#pragma unroll
        for (int k = 0; k < XPT; k++) x[b][k] = make_double2(0.0, 0.0);
      }
      if constexpr (KIN) wave_sum_to(ks, redk + wave * NB + b);
      if (b > 0) coriolis_energy(b - 1);
    }
    coriolis_energy(NB - 1);
    // first stage, per point: t_a = phi(l, i, a) sum_b V_l[a][b] x_b, in
    // place; on 256-point pencils with 6 to 9 blocks in the form that reads
    // each entry of V once for all points of the thread (on 9 blocks of
    // 256 x 192 x 96: 465 -> 383 us; on 192-point pencils it costs a wave per
    // SIMD and no time is gained)
    if constexpr (!kMixBatched) {
#pragma unroll
      for (int k = 0; k < XPT; k++) {
        const int e = ept(k);
        if (eok(k)) {
          const double2 *ph = (const double2 *)phs + (e % N1) * NB;
          double2 t[NB];
#pragma unroll
          for (int al = 0; al < NB; al++) {
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int b = 0; b < NB; b++) {
              const double v = V[min(al, nOc) * jb.nO + min(b, nOc)];
              acc.x = fma(v, x[b][k].x, acc.x);
              acc.y = fma(v, x[b][k].y, acc.y);
            }
            t[al] = cmul(ph[al], acc);
          }
#pragma unroll
          for (int al = 0; al < NB; al++) x[al][k] = t[al];
        }
      }
    } else {
      mix_first_stage<NB, XPT>(x, V, jb.nO, [&](int k, int al) {
        return ((const double2 *)phs)[(ept(k) % N1) * NB + al];
      });
    }
    // second stage, block by block: y_b = rot2(l, j) sum_a V_l[a][b] t_a,
    // stored; rotational energy l(l+1)(I1 + I2)|y|^2, |y|^2 and conj(c0) y.
    // Everything read from the job descriptor or the per-point tables is
    // loaded before the first store: as far as the compiler knows the stores
    // may alias the descriptor, and it would reload its fields (and then the
    // tables through them) after every store, two dependent round trips per
    // block.
    const double ll = (double)l * (l + 1);
    const int nO = jb.nO;
    const QD_G double2 *r2p = jb.rot2 + (long)l * jb.n2;
    QD_G double *pcor = jb.part_cor;
    const long ustride = jb.units[KK_FFT_RI];
    double2 rv[XPT];
    double iv[XPT];
#pragma unroll
    for (int k = 0; k < XPT; k++) {
      const int e = eok(k) ? ept(k) : 0;
      const int i = e % N1, j = g * P + e / N1;
      rv[k] = r2p[j];
      iv[k] = ll * (jb.I1[i] + jb.I2[j]);
    }
#pragma unroll
    for (int b = 0; b < NB; b++) {
      double sb[NF - 1];
#pragma unroll
      for (int f = 0; f < NF - 1; f++) sb[f] = 0.0;
      if (b < nb) {
        double2 z[ACF ? XPT : 1];
        if constexpr (ACF) {   // this block's c0 values ahead of its stores
#pragma unroll
          for (int k = 0; k < XPT; k++) z[k] = c0[b * LM + (eok(k) ? ept(k) : 0)];
        }
        // sum_a V[a][b] t_a for every point (batched: each entry of V read once)
        double2 accs[kMixBatched ? XPT : 1];
        if constexpr (kMixBatched) {
#pragma unroll
          for (int k = 0; k < XPT; k++) accs[k] = make_double2(0.0, 0.0);
#pragma unroll
          for (int al = 0; al < NB; al++) {
            const double v = V[min(al, nOc) * nO + b];
#pragma unroll
            for (int k = 0; k < XPT; k++) {
              accs[k].x = fma(v, x[al][k].x, accs[k].x);
              accs[k].y = fma(v, x[al][k].y, accs[k].y);
            }
          }
        }
#pragma unroll
        for (int k = 0; k < XPT; k++) {
          const int e = ept(k);
          if (eok(k)) {
            double2 acc = make_double2(0.0, 0.0);
            if constexpr (kMixBatched) {
              acc = accs[k];
            } else {
#pragma unroll
              for (int al = 0; al < NB; al++) {
                const double v = V[min(al, nOc) * nO + b];
                acc.x = fma(v, x[al][k].x, acc.x);
                acc.y = fma(v, x[al][k].y, acc.y);
              }
            }
            const double2 y = cmul(rv[k], acc);
            store_out(base + b * LM + e, y, stream);
            const double q = cnorm(y);
            sb[0] += iv[k] * q;
            sb[1] += q;
            if constexpr (ACF) {
              sb[2] += z[k].x * y.x + z[k].y * y.y;
              sb[3] += z[k].x * y.y - z[k].y * y.x;
            }
          }
        }
      }
#pragma unroll
      for (int f = 0; f < NF - 1; f++) wave_sum_to(sb[f], red + wave * NV + b * NF + 1 + f);
    }
    __syncthreads();
    if (threadIdx.x < NF * nO) {
      const int q = threadIdx.x;   // q = NF b + f, stored at field f of block b ([b][5][units])
      const double sum = (q < NV) ? (red[q] + red[NV + q]) + (red[2 * NV + q] + red[3 * NV + q]) : 0.0;
      pcor[(long)(5 * (q / NF) + q % NF) * ustride + u] = sum;
    }
    if constexpr (KIN) {
      // kinetic energy of block b in this workgroup's rows, as the 2-D
      // momentum-space sum sum |F psi|^2 (E1 + E2) that the reduction scales.
      // R part: the sums above over n2 (the r pass leaves its output
      // multiplied by n2, so they are n2^2 sum |F_R psi|^2 E1, and by
      // Parseval along r the 2-D sum is n2 times that; the r factor is
      // unitary along r and changes neither).  r part: n1 times the r pass's
      // sums (Parseval along R), those of the slab's column groups g,
      // g + groups, .. taken here in that order.
      if ((int)threadIdx.x < nO) {
        const int b = threadIdx.x;
        const double e1 = (redk[b] + redk[NB + b]) + (redk[2 * NB + b] + redk[3 * NB + b]);
        const int gcols = jb.n1 / jb.mix_cols;
        const QD_G double *k2 =
            jb.part_ke2 + (long)b * jb.units[KK_FFT_r] + (long)(l - jb.omin) * gcols;
        double e2 = 0.0;
        for (int ga = g; ga < gcols; ga += groups) e2 += k2[ga];
        jb.part_ke[(long)b * ustride + u] = e1 / (double)jb.n2 + (double)N1 * e2;
      }
    }
    // fused Legendre step: this step's backward transform is now pending
    if (u == 0 && threadIdx.x == 0 && jb.leg_step) jb.counters[3] = 1;
  }
}

// ---------------------------------------------------------------------------
// r pass: forward DFT along r, kinetic phase and energy, inverse DFT.
// PlanEnds makes the boundary stages of both transforms take the points
// (p, idx + q n2/R1) of butterfly (p, idx): the forward first stage reads
// them from HBM, the forward last stage ends with them in registers, where
// the kinetic phase is applied and the inverse first stage runs at once, and
// the inverse last stage writes them to HBM.  LDS sees only the middle
// stages: 8 accesses per point instead of 20 with an LDS round trip per stage.
// ---------------------------------------------------------------------------
template <int N2>
__global__ __launch_bounds__(kBlock) void k_fft_r_ct(StepArgs a) {
  constexpr int Q = ct_cols(N2);
  constexpr int QN = Q * N2;
  constexpr int R1 = PlanEnds::first(N2), NB1 = N2 / R1;   // radix and butterflies per pencil at both ends
  constexpr int NSL = N2 / R1;                           // stride of the last stage (= NB1)
  constexpr int T1 = Q * NB1, MB = (T1 + kBlock - 1) / kBlock;
  static_assert(N2 % (R1 * R1) == 0, "PlanEnds needs n2 to be a multiple of R1^2");
  __shared__ double2 A[QN];   // pencil p element j at A[p + j Q]
  __shared__ double2 tw[N2];
  __shared__ double red[4];
  const int job = locate_job(a.unit_base, a.n_jobs, blockIdx.x);
  const JobDev &jb = a.jobs[job];
  const int stream = jb.stream_out;   // read before any store (stores may alias the descriptor)
  const int u = blockIdx.x - a.unit_base[job];
  const int n1 = jb.n1;
  int b, l, g;
  decode_slab_unit(jb, u, n1 / Q, b, l, g);
  QD_G double2 *base = jb.fbr + ((long)b * jb.L + l) * jb.M + (long)g * Q;   // (p, j) at base[j n1 + p]
  // twiddles loaded here and written to LDS after the pencil loads below, so
  // that both round trips overlap
  constexpr int NTW = (N2 + kBlock - 1) / kBlock;
  double2 tv[NTW];
#pragma unroll
  for (int it = 0; it < NTW; it++) tv[it] = jb.twr[min((int)threadIdx.x + it * kBlock, N2 - 1)];
  double2 X[MB][R1];
  // forward first stage (stride 1) on points read from HBM
#pragma unroll
  for (int m = 0; m < MB; m++) {
    const int t = threadIdx.x + m * kBlock;
    if (T1 % kBlock == 0 || t < T1) {
      const int p = t % Q, idx = t / Q;
      double2 v[R1];
#pragma unroll
      for (int q = 0; q < R1; q++) v[q] = base[(long)(idx + q * NB1) * n1 + p];
      dft_small<false, R1>(v, X[m]);
#pragma unroll
      for (int q = 0; q < R1; q++) A[p + (idx * R1 + q) * Q] = X[m][q];
    }
  }
#pragma unroll
  for (int it = 0; it < NTW; it++)
    if ((int)threadIdx.x + it * kBlock < N2) tw[threadIdx.x + it * kBlock] = tv[it];
  __syncthreads();
  stages_ip<false, N2, R1, Q, 1, Q, true, PlanEnds, NSL>(A, tw);
  // forward last stage (stride NSL, twiddles w^(q idx)) into registers:
  // outputs at j = idx + q NSL; kinetic phase and energy; inverse first stage
  double ke = 0.0;
#pragma unroll
  for (int m = 0; m < MB; m++) {
    const int t = threadIdx.x + m * kBlock;
    if (T1 % kBlock == 0 || t < T1) {
      const int p = t % Q, idx = t / Q;
      const int i = g * Q + p;
      double2 v[R1], y[R1];
#pragma unroll
      for (int q = 0; q < R1; q++) v[q] = A[p + (idx + q * NB1) * Q];
      if constexpr (NSL > 1) {
#pragma unroll
        for (int q = 1; q < R1; q++) v[q] = cmul(v[q], tw[q * idx]);
      }
      dft_small<false, R1>(v, y);
      const double2 k1 = jb.kin1[i];
      const double e1 = jb.kinE1[i];
#pragma unroll
      for (int q = 0; q < R1; q++) {
        const int j = idx + q * NSL;
        const double2 x = cmul(y[q], cmul(k1, jb.kin2[j]));
        ke += cnorm(x) * (e1 + jb.kinE2[j]);
        y[q] = x;
      }
      dft_small<true, R1>(y, X[m]);
    }
  }
  __syncthreads();   // every read of the last stage precedes the in-place writes
#pragma unroll
  for (int m = 0; m < MB; m++) {
    const int t = threadIdx.x + m * kBlock;
    if (T1 % kBlock == 0 || t < T1) {
      const int p = t % Q, idx = t / Q;
#pragma unroll
      for (int q = 0; q < R1; q++) A[p + (idx * R1 + q) * Q] = X[m][q];
    }
  }
  __syncthreads();
  stages_ip<true, N2, R1, Q, 1, Q, true, PlanEnds, NSL>(A, tw);
  // inverse last stage, written to HBM
#pragma unroll
  for (int m = 0; m < MB; m++) {
    const int t = threadIdx.x + m * kBlock;
    if (T1 % kBlock == 0 || t < T1) {
      const int p = t % Q, idx = t / Q;
      double2 v[R1], y[R1];
#pragma unroll
      for (int q = 0; q < R1; q++) v[q] = A[p + (idx + q * NB1) * Q];
      if constexpr (NSL > 1) {
#pragma unroll
        for (int q = 1; q < R1; q++) {
          double2 w = tw[q * idx];
          w.y = -w.y;
          v[q] = cmul(v[q], w);
        }
      }
      dft_small<true, R1>(v, y);
#pragma unroll
      for (int q = 0; q < R1; q++) store_out(base + (long)(idx + q * NSL) * n1 + p, y[q], stream);
    }
  }
  ke = wave_sum64(ke);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ke;
  __syncthreads();
  if (threadIdx.x == 0) jb.part_ke[u] = (red[0] + red[1]) + (red[2] + red[3]);
}

// ---------------------------------------------------------------------------
// Two-pass form, r pass (where mix_cols(n2, blocks) > 0).  The step's kinetic
// factor F^-1 exp(-i dt (E1 + E2)) F, F = F_r F_R the 2-D transform, is the
// product of the R factor F_R^-1 kin1 F_R and the r factor F_r^-1 kin2 F_r
// (they act on different indices).  This pass applies the first rotational .
// Coriolis . rotational half step (the forward R pass's per-point mix) and
// the r factor; the inverse R pass (KIN) applies the R factor and the second
// half step: two read-write passes over the FBR array per step instead of
// three.  The kinetic energy of the state after the first half step
// (cudaOpenMP2.cu:358-360) splits the same way: its r part is summed here
// (Parseval along R), its R part in the R pass.
// Workgroup (l, g) takes columns i = g Q .. g Q + Q - 1 of slab l in every
// block present; each block's points stay in registers from the load to its
// first DFT stage, as the inputs of the thread's butterflies of radix
// mix_radix.  The output is left multiplied by n2 (unscaled transform pair):
// the R pass scales by 1 / (n1 n2).  At most 3 waves per SIMD (the register
// budget of mix_cols; 4 would spill), and the two-transform R pass of 5
// blocks of 192 points at 4 (4 registers spill; on 192 x 128 x 64 with 5
// blocks 77 against 82 us at 3, measured before its blocks were loaded two
// ahead).
// ---------------------------------------------------------------------------
template <int N2, int NB>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_fft_r_mix(StepArgs a) {
  constexpr int Q = mix_cols(N2, NB);
  static_assert(Q > 0 && kBlock % Q == 0, "no r-pass tiling with the mix for these sizes");
  constexpr int QN = Q * N2;
  using RP = std::conditional_t<mix_radix(N2, NB) == 8, PlanEnds, PlanEnds4>;
  constexpr int R1 = RP::first(N2), NB1 = N2 / R1, NSL = NB1;
  static_assert(R1 == mix_radix(N2, NB) && N2 % (R1 * R1) == 0, "end stages of radix R1 need n2 % R1^2 == 0");
  constexpr int T1 = Q * NB1, MB = (T1 + kBlock - 1) / kBlock;
  __shared__ double2 A[QN];   // column p element j at A[p + j Q]
  __shared__ double2 tw[N2];
  __shared__ double red[4 * NB];
  if (reduce_carried(a, red)) return;   // leading workgroups: the previous step's reduction
  if (flux_carried(a)) return;          // then its flux surface values
  const int bid = blockIdx.x - a.red_units - a.flux_units;
  const int job = locate_job(a.unit_base, a.n_jobs, bid);
  const JobDev &jb = a.jobs[job];
  const int stream = jb.stream_out;   // read before any store (stores may alias the descriptor)
  const int u = bid - a.unit_base[job];
  const int n1 = jb.n1, nO = jb.nO;
  const long LM = (long)jb.L * jb.M;
  const int groups = n1 / Q;
  const int l = jb.omin + u / groups, g = u % groups;
  const int nb = min(jb.kmax, l) - jb.omin + 1;
  const int p = threadIdx.x % Q, i = g * Q + p;   // the thread's column (Q divides kBlock)
  QD_G double2 *base = jb.fbr + (long)l * jb.M + (long)g * Q;   // block b at + b LM; (p, j) at [j n1 + p]
  QD_G double *pke = jb.part_ke2 + u;
  const long ua = jb.units[KK_FFT_r];
  auto live = [](int m) { return T1 % kBlock == 0 || (int)threadIdx.x + m * kBlock < T1; };
  auto bfly = [&](int m) { return live(m) ? ((int)threadIdx.x + m * kBlock) / Q : 0; };
  // every block's points of the thread's butterflies (inputs j = idx + q NB1
  // of the first stage), in flight at once; blocks >= nb load block nb - 1
  // (values never used), idle threads butterfly 0
  double2 x[NB][MB][R1];
#pragma unroll
  for (int b = 0; b < NB; b++)
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
      for (int q = 0; q < R1; q++)
        x[b][m][q] = base[(b < nb ? b : nb - 1) * LM + (long)(bfly(m) + q * NB1) * n1 + p];
  constexpr int NTW = (N2 + kBlock - 1) / kBlock;
  double2 tv[NTW];
#pragma unroll
  for (int it = 0; it < NTW; it++) tv[it] = jb.twr[min((int)threadIdx.x + it * kBlock, N2 - 1)];
  double2 ph[NB];   // corPhi(l, i, a)
#pragma unroll
  for (int al = 0; al < NB; al++) ph[al] = jb.corPhi[((long)l * n1 + i) * nO + min(al, nO - 1)];
#pragma unroll
  for (int it = 0; it < NTW; it++)
    if ((int)threadIdx.x + it * kBlock < N2) tw[threadIdx.x + it * kBlock] = tv[it];
  // first half step, point by point (V entries scalar, as in the forward R pass)
  {
    const QD_C double *V = (const QD_C double *)(jb.corV + (long)l * nO * nO);
    FbrSums<NB, 3> none;
#pragma unroll
    for (int m = 0; m < MB; m++) {
      if (live(m)) {
#pragma unroll
        for (int q = 0; q < R1; q++) {
          double2 xp[NB];
#pragma unroll
          for (int b = 0; b < NB; b++) xp[b] = x[b][m][q];
          coriolis_point<false, NB, 3>(jb, xp, nb, l, i, bfly(m) + q * NB1, V, none, nullptr, LM,
                                       (const double2 *)ph);
#pragma unroll
          for (int b = 0; b < NB; b++) x[b][m][q] = xp[b];
        }
      }
    }
  }
  // the r factor and energy at the last stage's outputs j = idx + q NSL, the
  // same for every block: read before the first store
  double2 kf[MB][R1];
  double ef[MB][R1];
#pragma unroll
  for (int m = 0; m < MB; m++)
#pragma unroll
    for (int q = 0; q < R1; q++) {
      kf[m][q] = jb.kin2[bfly(m) + q * NSL];
      ef[m][q] = jb.kinE2[bfly(m) + q * NSL];
    }
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int b = 0; b < NB; b++) {
    double ke = 0.0;   // r kinetic energy of the thread's points of block b
    if (b < nb) {
      // the twiddles (first block) or the previous block's last-stage reads
      // before these writes
      __syncthreads();
      // forward first stage (stride 1) from registers
#pragma unroll
      for (int m = 0; m < MB; m++) {
        if (live(m)) {
          double2 X[R1];
          dft_small<false, R1>(x[b][m], X);
#pragma unroll
          for (int q = 0; q < R1; q++) A[p + (bfly(m) * R1 + q) * Q] = X[q];
        }
      }
      __syncthreads();
      stages_ip<false, N2, R1, Q, 1, Q, true, RP, NSL>(A, tw);
      // forward last stage into registers, the r factor and energy, the
      // inverse first stage
#pragma unroll
      for (int m = 0; m < MB; m++) {
        if (live(m)) {
          const int idx = bfly(m);
          double2 v[R1], y[R1];
#pragma unroll
          for (int q = 0; q < R1; q++) v[q] = A[p + (idx + q * NB1) * Q];
          if constexpr (NSL > 1) {
#pragma unroll
            for (int q = 1; q < R1; q++) v[q] = cmul(v[q], tw[q * idx]);
          }
          dft_small<false, R1>(v, y);
#pragma unroll
          for (int q = 0; q < R1; q++) {
            ke += cnorm(y[q]) * ef[m][q];
            y[q] = cmul(y[q], kf[m][q]);
          }
          dft_small<true, R1>(y, x[b][m]);
        }
      }
      __syncthreads();   // every read of the last stage precedes the in-place writes
#pragma unroll
      for (int m = 0; m < MB; m++) {
        if (live(m)) {
#pragma unroll
          for (int q = 0; q < R1; q++) A[p + (bfly(m) * R1 + q) * Q] = x[b][m][q];
        }
      }
      __syncthreads();
      stages_ip<true, N2, R1, Q, 1, Q, true, RP, NSL>(A, tw);
      // inverse last stage, written to HBM
#pragma unroll
      for (int m = 0; m < MB; m++) {
        if (live(m)) {
          const int idx = bfly(m);
          double2 v[R1], y[R1];
#pragma unroll
          for (int q = 0; q < R1; q++) v[q] = A[p + (idx + q * NB1) * Q];
          if constexpr (NSL > 1) {
#pragma unroll
            for (int q = 1; q < R1; q++) {
              double2 w = tw[q * idx];
              w.y = -w.y;
              v[q] = cmul(v[q], w);
            }
          }
          dft_small<true, R1>(v, y);
#pragma unroll
          for (int q = 0; q < R1; q++) store_out(base + b * LM + (long)(idx + q * NSL) * n1 + p, y[q], stream);
        }
      }
    }
    wave_sum_to(ke, red + wave * NB + b);   // 0 for blocks absent at slab l
  }
  // r kinetic sums of the workgroup per block, [b][units]
  __syncthreads();
  if ((int)threadIdx.x < nO) {
    const int b = threadIdx.x;
    pke[b * ua] = (red[b] + red[NB + b]) + (red[2 * NB + b] + red[3 * NB + b]);
  }
}

// ---------------------------------------------------------------------------
// launchers for one compile-time R length (instantiated in kernels_fft_ct*.hip)
// ---------------------------------------------------------------------------
template <bool INV, int N1, bool ACF>
void launch_R_nb(const StepArgs &a, int units, int nb_max, hipStream_t s) {
  const dim3 g(units + (INV ? 0 : a.red_units + a.flux_units)), t(kBlock);
  if (nb_max <= 1) hipLaunchKernelGGL((k_fft_R_ct<INV, 1, N1, ACF>), g, t, 0, s, a);
  else if (nb_max <= 2) hipLaunchKernelGGL((k_fft_R_ct<INV, 2, N1, ACF>), g, t, 0, s, a);
  else if (nb_max <= 3) hipLaunchKernelGGL((k_fft_R_ct<INV, 3, N1, ACF>), g, t, 0, s, a);
  else if (nb_max <= 5) hipLaunchKernelGGL((k_fft_R_ct<INV, 5, N1, ACF>), g, t, 0, s, a);
  else if (nb_max <= 7) hipLaunchKernelGGL((k_fft_R_ct<INV, 7, N1, ACF>), g, t, 0, s, a);
  else if (nb_max <= 9) hipLaunchKernelGGL((k_fft_R_ct<INV, 9, N1, ACF>), g, t, 0, s, a);
  else if (nb_max <= 16) hipLaunchKernelGGL((k_fft_R_ct<INV, 16, N1, ACF>), g, t, 0, s, a);
  else hipLaunchKernelGGL((k_fft_R_ct<INV, 24, N1, ACF>), g, t, 0, s, a);
}

// two-pass form: the inverse R pass with the R kinetic factor (up to 5 blocks)
template <int N1, bool ACF>
void launch_R_kin_nb(const StepArgs &a, int units, int nb_max, hipStream_t s) {
  const dim3 g(units), t(kBlock);
  if (nb_max <= 1) hipLaunchKernelGGL((k_fft_R_ct<true, 1, N1, ACF, true>), g, t, 0, s, a);
  else if (nb_max <= 2) hipLaunchKernelGGL((k_fft_R_ct<true, 2, N1, ACF, true>), g, t, 0, s, a);
  else if (nb_max <= 3) hipLaunchKernelGGL((k_fft_R_ct<true, 3, N1, ACF, true>), g, t, 0, s, a);
  else hipLaunchKernelGGL((k_fft_R_ct<true, 5, N1, ACF, true>), g, t, 0, s, a);
}
// and the r pass with the first half step, for the block counts it has a tiling for
template <int N2, int NB>
void launch_r_mix_one(const StepArgs &a, dim3 g, hipStream_t s) {
  if constexpr (mix_cols(N2, NB) > 0) hipLaunchKernelGGL((k_fft_r_mix<N2, NB>), g, dim3(kBlock), 0, s, a);
}
template <int N2>
void launch_r_mix_nb(const StepArgs &a, int units, int nb_max, hipStream_t s) {
  const dim3 g(units + a.red_units + a.flux_units);
  if (nb_max <= 1) launch_r_mix_one<N2, 1>(a, g, s);
  else if (nb_max <= 2) launch_r_mix_one<N2, 2>(a, g, s);
  else if (nb_max <= 3) launch_r_mix_one<N2, 3>(a, g, s);
  else launch_r_mix_one<N2, 5>(a, g, s);
}

#define QD_FFT_CT_DEFINE(N1)                                                                   \
  void launch_fft_R_ct_##N1(const StepArgs &a, int units, int inverse, int acf, int nb_max, hipStream_t s) { \
    if (!inverse) launch_R_nb<false, N1, false>(a, units, nb_max, s);                          \
    else if (acf) launch_R_nb<true, N1, true>(a, units, nb_max, s);                            \
    else launch_R_nb<true, N1, false>(a, units, nb_max, s);                                    \
  }                                                                                            \
  void launch_fft_r_ct_##N1(const StepArgs &a, int units, hipStream_t s) {                     \
    hipLaunchKernelGGL(k_fft_r_ct<N1>, dim3(units), dim3(kBlock), 0, s, a);                    \
  }                                                                                            \
  void launch_fft_R_kin_##N1(const StepArgs &a, int units, int acf, int nb_max, hipStream_t s) { \
    if (acf) launch_R_kin_nb<N1, true>(a, units, nb_max, s);                                   \
    else launch_R_kin_nb<N1, false>(a, units, nb_max, s);                                      \
  }                                                                                            \
  void launch_fft_r_mix_##N1(const StepArgs &a, int units, int nb_max, hipStream_t s) {        \
    launch_r_mix_nb<N1>(a, units, nb_max, s);                                                  \
  }

}  // namespace qd
