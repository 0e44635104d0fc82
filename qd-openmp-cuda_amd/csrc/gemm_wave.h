// gemm_wave.h - per-wave building blocks of the Legendre GEMMs on the matrix
// cores: operand loads, B staging into LDS, the FP64 and mixed-precision
// multiply-accumulate of a strip, and the per-wave strip loops of the
// forward and backward passes (kernels_gemm.hip, kernels_leg_step.hip).
#pragma once

#include <cstdlib>

#include "qd_internal.h"
#include "qd_complex.h"

namespace qd {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// bfloat16 operand of the compensation MFMA: slots 0..3 the hi parts, 4..7
// the lo parts, each an FP32 value rounded to nearest even
// (v_cvt_pk_bf16_f32, two values per instruction).  The empty asm operand
// keeps the compiler from folding the conversion with the FP64 -> FP32
// rounding that produced the value into a direct FP64 -> bfloat16
// conversion, which it would emulate with compares and selects.
__device__ __forceinline__ __bf16 bf16_of(float x) {
  asm("" : "+v"(x));
  return (__bf16)x;
}
__device__ __forceinline__ v8bf bf16_operand(const float (&h)[4], const float (&l)[4]) {
  return v8bf{bf16_of(h[0]), bf16_of(h[1]), bf16_of(h[2]), bf16_of(h[3]),
              bf16_of(l[0]), bf16_of(l[1]), bf16_of(l[2]), bf16_of(l[3])};
}

__host__ __device__ constexpr int lds_stride(int nt_tiles) { return nt_tiles * 16 + ((nt_tiles & 1) ? 0 : 16); }

// k-steps (of 4 k each) per pipelined tile: a template parameter KC in
// {8, 12, 16}; KC = 8 runs three tiles deep (up to 6 column tiles; the mixed
// form up to 4), 12 and 16 two deep.  Deeper
// pipelines keep more bytes in flight per wave but cost registers; the choice
// keeps two waves per SIMD (VGPRs + AGPRs <= 256) where the tile allows it.
// The forward pass also holds the phase operands and runs KC = 8 three deep;
// the backward pass holds the potentials of the epilogue and runs one or two
// wide tiles per strip, except in the mixed form, whose FP32 hi/lo operands
// and partial accumulators leave two waves per SIMD only at KC = 8.
inline int leg_chunk(int ksteps, bool fwd, int nt_tiles, bool mixed) {
  if (fwd || nt_tiles > 4 || mixed) return 8;
  int best = 8, waste = 1 << 30;
  for (int kc : {16, 12, 8}) {
    const int w = (ksteps + kc - 1) / kc * kc - ksteps;
    if (w < waste) { waste = w; best = kc; }
  }
  return best;
}

// Logical workgroup index with the workgroups of one XCD contiguous: the
// hardware deals workgroups round-robin over the 8 XCDs (XCD = id % 8), each
// with its own L2; consecutive logical units (the omega blocks of one row
// group, which read the same phase / potential rows) then share an L2.
__device__ __forceinline__ int xcd_logical_unit(int p, int total) {
  const int q = total >> 3, r = total & 7, x = p & 7, slot = p >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + slot;
}

// One wave's view of a Legendre GEMM: C (rows x N, ldc) = op(A) (rows x K, lda) . B.
struct GemmWave {
  const double *A;
  long lda;
  int K, ksteps;
  double *C;
  long ldc;
  int N;
  long rows;
  const double *Bs;              // FP64 B in LDS, [kp][Ws]
  const float *Bh, *Bl;          // FP32 hi/lo B in LDS (mixed form)
  const v4u *Bx;                 // paired mixed form: bfloat16 (bl, bh) per chunk, lane group and column
  int Ws;
  const QD_G double *vph;        // forward: phase table viewed as real (2M x nt), same indexing as A
  const QD_G double *wpot;       // backward epilogue: w_k V(k, m) [NT * 16][M], zero rows k >= N
  long M;
};

// Raw buffer access (32-bit offsets, hardware range check): a load whose byte
// offset is at or past `bytes` returns 0.  The operand tiles read rows k >= K
// of the k-padded product this way, so the zero padding costs no instruction.
// Dword 3 holds the gfx9-family buffer format word.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double buffer_load_f64(__amdgpu_buffer_rsrc_t r, uint32_t voff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, 0, 0));
}

// B ([kp][np] row-major, global) into LDS rows of stride Ws, columns np..W-1
// zeroed.  Each thread issues up to 8 loads before its first LDS write, so the
// staging costs one round trip per 8 elements per thread rather than one per
// element.
template <int TB, int W, typename T, typename Src>
__device__ __forceinline__ void stage_b(T *Bs, Src Bg, int kp, int np, int Ws) {
  constexpr int CH = 8;
  const int total = kp * W;
  for (int e0 = 0; e0 < total; e0 += CH * TB) {
    T v[CH];
#pragma unroll
    for (int q = 0; q < CH; q++) {
      const int e = e0 + q * TB + (int)threadIdx.x;
      const int k = e / W, c = e - k * W;
      v[q] = (e < total && c < np) ? (T)Bg[k * np + c] : (T)0;
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

// Value of the neighbouring lane (lane ^ 1) through a DPP quad permutation
// (1, 0, 3, 2): a VALU move, no LDS round trip.
__device__ __forceinline__ double swap_adjacent(double x) {
  const long long v = __double_as_longlong(x);
  int lo = (int)(v & 0xffffffffLL), hi = (int)(v >> 32);
  lo = __builtin_amdgcn_mov_dpp(lo, 0xB1, 0xF, 0xF, false);
  hi = __builtin_amdgcn_mov_dpp(hi, 0xB1, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Column of accumulator item `it` of tile t for this lane.
// v_mfma_f64_16x16x4f64: item it of lane (lr, lk) holds D[lk + 4 it][lr];
// v_mfma_f32_16x16x4f32: item it holds D[4 lk + it][lr] (D = C^T tile).
template <bool MIXED>
__device__ __forceinline__ int acc_col(int t, int lk, int it) {
  return MIXED ? t * 16 + 4 * lk + it : t * 16 + lk + 4 * it;
}

// Computes the strips strip0 + s * sstride (s < nstrips) of one wave.  The
// (strip, k-chunk) tiles are software-pipelined DEPTH deep: the loads of tile
// t + DEPTH - 1 (kCh A values and, forward, kCh phase components per lane)
// are issued before tile t runs on the matrix cores, into a ring of DEPTH
// register buffers indexed at compile time (the tile loop is unrolled by
// DEPTH).  Loads are unconditional from clamped, always in-bounds addresses
// (tiles past the end reload the last one) and raw values are masked only
// where they are used, so the compiler waits for exactly the oldest tile.
// The backward epilogue's potential loads are issued before the prefetch of
// the same tile, so waiting for them does not wait for the prefetch.
//   PHASE: multiply each A element by d1 d2 exp(-i dt V) (forward prologue);
//          a lane holds one real component, its partner lane (r ^ 1) the other.
//   OBS:   accumulate w V |psi|^2 of the output (backward epilogue) into ob.
// `stage` (the workgroup's staging of B into LDS and its barrier) runs once on
// every path, after the first tiles' loads are issued, so that both round
// trips overlap.
template <int NT, int kCh, int DEPTH, bool MIXED, bool PHASE, bool OBS, typename Stage>
__device__ __forceinline__ void gemm_wave(const GemmWave &gw, long strip0, int nstrips, int sstride, double &ob,
                                          Stage &&stage) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const bool odd = lr & 1;
  const int nch = (gw.ksteps + kCh - 1) / kCh;
  int ns = 0;
  for (int s = 0; s < nstrips; s++)
    if ((strip0 + (long)s * sstride) * 16 < gw.rows) ns = s + 1;
  const int ntiles = ns * nch;
  if (ntiles <= 0) {
    stage();
    return;
  }
  double abuf[DEPTH][kCh], tbuf[DEPTH][kCh];
  auto row_of = [&](int tile) { return (strip0 + (long)(tile / nch) * sstride) * 16 + lr; };
  // operand rows k < K only: rows K.. of the padded product read as zero
  const uint32_t abytes = (uint32_t)((unsigned long)gw.K * (unsigned long)gw.lda * 8ul);
  const __amdgpu_buffer_rsrc_t rA = make_rsrc(gw.A, abytes);
  const __amdgpu_buffer_rsrc_t rT = make_rsrc(PHASE ? (const void *)gw.vph : (const void *)gw.A, abytes);
  const uint32_t kstride = (uint32_t)gw.lda * 32u;   // bytes between k and k + 4
  // The forward pass (two operands per element) loads through buffer
  // descriptors; the backward pass keeps clamped global loads, which leave it
  // the registers for two waves per SIMD, and masks rows k >= K where they
  // are used.
  constexpr bool BUF = PHASE;
  auto load = [&](int tile, double(&a)[kCh], double(&tp)[kCh]) {
    tile = min(tile, ntiles - 1);
    const int c = tile % nch;
    const long r = row_of(tile);
    // rows past the end reread the last row: their products land in output
    // rows that are never stored (row r of C depends on row r of A only)
    const long rc = r < gw.rows ? r : gw.rows - 1;
    if (BUF) {
      const uint32_t v0 = ((uint32_t)(4 * c * kCh + lk) * (uint32_t)gw.lda + (uint32_t)rc) * 8u;
#pragma unroll
      for (int q = 0; q < kCh; q++) {
        a[q] = buffer_load_f64(rA, v0 + q * kstride);
        if (PHASE) tp[q] = buffer_load_f64(rT, v0 + q * kstride);   // even lane Re T, odd lane Im T
      }
    } else {
      // a uniform row base per k plus one 32-bit lane offset (the compiler
      // still forms most addresses as 64-bit vector pairs).
      // k-steps past the last (when kCh does not divide ksteps) reread the
      // last one; rows K <= k < 4 ksteps are allocated (zero padding past
      // the last block) and masked where they are used.
      const uint32_t vb = ((uint32_t)lk * (uint32_t)gw.lda + (uint32_t)rc) * 8u;
#pragma unroll
      for (int q = 0; q < kCh; q++) {
        const int kq = min(c * kCh + q, gw.ksteps - 1);
        const char *Aq = (const char *)(gw.A + (long)(4 * kq) * gw.lda);
        a[q] = *(const double *)(Aq + vb);
      }
    }
  };
  // FP64 sums (acc); mixed: the hi*hi sums of every 4 k-steps (16 k) in
  // FP32 are added into acc at once, the compensation terms (xx, 2^-24 of
  // the total) accumulate in FP32 over the strip and join in the epilogue
  v4d acc[NT];
  v4f xx[MIXED ? NT : 1];
  auto reset = [&]() {
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
    if constexpr (MIXED) {
#pragma unroll
      for (int t = 0; t < NT; t++) xx[t] = v4f{0.f, 0.f, 0.f, 0.f};
    }
  };
  reset();
  auto compute = [&](int tile, const double(&an)[kCh], const double(&tn)[kCh], const double(&vh)[NT][2]) {
    const int c = tile % nch;
    const long r = row_of(tile);
    const bool rok = r < gw.rows;
    if (c == 0) reset();
    double ac[kCh];
#pragma unroll
    for (int q = 0; q < kCh; q++) {
      const double av = (BUF || (rok && 4 * (c * kCh + q) + lk < gw.K)) ? an[q] : 0.0;
      if (PHASE) {
        const double tv = tn[q];
        const double oa = swap_adjacent(av);
        const double ot = swap_adjacent(tv);
        const double Tr = odd ? ot : tv, Ti = odd ? tv : ot;
        // re' = re Tr - im Ti (even lane), im' = re Ti + im Tr (odd lane)
        ac[q] = fma(odd ? oa : -oa, Ti, av * Tr);
      } else {
        ac[q] = av;
      }
    }
    if constexpr (!MIXED) {
#pragma unroll
      for (int q = 0; q < kCh; q++) {
        if (c * kCh + q < gw.ksteps) {
          const double *brow = gw.Bs + (4 * (c * kCh + q) + lk) * gw.Ws + lr;
#pragma unroll
          for (int t = 0; t < NT; t++)
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(brow[t * 16], ac[q], acc[t], 0, 0, 0);
        }
      }
    } else {
      // a = ah + al, B = bh + bl in FP32 hi and lo parts: hi*hi on the FP32
      // MFMA per k-step; hi*lo + lo*hi (2^-24 of the product) on one bf16
      // MFMA per 4 k-steps (16 k), whose 8 operand slots pair ah(k_q) with
      // bl(k_q) and al(k_q) with bh(k_q) for the lane's k-rows
      // k_q = 16 c16 + lk + 4 q (the table gw.Bx; see paired_mac)
      static_assert(kCh % 4 == 0, "k-chunks of whole compensation groups");
#pragma unroll
      for (int q0 = 0; q0 < kCh; q0 += 4) {
        if (c * kCh + q0 < gw.ksteps) {
          float hb[4], lb[4];
          v4f hh[NT];
#pragma unroll
          for (int t = 0; t < NT; t++) hh[t] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int q = q0; q < q0 + 4; q++) {
            const float ah = (float)ac[q];
            hb[q - q0] = ah;
            lb[q - q0] = (float)(ac[q] - (double)ah);
            if (c * kCh + q < gw.ksteps) {
              const float *brow = gw.Bh + (4 * (c * kCh + q) + lk) * gw.Ws + lr;
#pragma unroll
              for (int t = 0; t < NT; t++)
                hh[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(brow[t * 16], ah, hh[t], 0, 0, 0);
            }
          }
#pragma unroll
          for (int t = 0; t < NT; t++)
#pragma unroll
            for (int it = 0; it < 4; it++) acc[t][it] += (double)hh[t][it];
          const v8bf xa = bf16_operand(hb, lb);
          const v4u *xrow = gw.Bx + ((c * kCh + q0) + lk) * (NT * 16) + lr;
#pragma unroll
          for (int t = 0; t < NT; t++)
            xx[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, xrow[t * 16]), xa, xx[t], 0, 0, 0);
        }
      }
    }
    if (c == nch - 1) {   // epilogue of the strip
      // w V of item it: this lane loaded items 2 odd + {0, 1}, its partner
      // lane (the other component of the same point) the other two
      double vo[NT][2];
      if (OBS) {
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
          for (int h = 0; h < 2; h++) vo[t][h] = swap_adjacent(vh[t][h]);
      }
#pragma unroll
      for (int t = 0; t < NT; t++) {
#pragma unroll
        for (int it = 0; it < 4; it++) {
          const int col = acc_col<MIXED>(t, lk, it);
          const bool ok = rok && col < gw.N;
          double v = acc[t][it];
          if constexpr (MIXED) v += (double)xx[t][it];
          // streaming store: the output is read once, by the next pass; kept
          // out of L2 it leaves room for the reused tables (backward GEMM on
          // the J=20 workload: 74 -> 66 us)
          if (ok) __builtin_nontemporal_store(v, &gw.C[(long)col * gw.ldc + r]);
          if (OBS) {
            const double wv = ((it >> 1) == (int)odd) ? vh[t][it & 1] : vo[t][it & 1];   // 0 for col >= N
            ob += (rok ? v * v : 0.0) * wv;
          }
        }
      }
    }
  };
  // weighted potentials w_k V of the strip of `tile` (backward epilogue),
  // issued before the prefetch.  The two lanes of one point (rows 2m, 2m + 1)
  // need the same 4 NT values: each loads the half with items it = 2 odd + h
  // and the epilogue swaps the halves (the table holds NT * 16 theta rows,
  // zero past N: no clamp).
  auto load_pot = [&](int tile, double(&vh)[NT][2]) {
    const long r = row_of(tile);
    const uint32_t m = (uint32_t)((r < gw.rows ? r : gw.rows - 1) >> 1);
    const uint32_t vb = ((uint32_t)acc_col<MIXED>(0, lk, 2 * (int)odd) * (uint32_t)gw.M + m) * 8u;
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int c0 = acc_col<MIXED>(t, 0, h);   // uniform part of the column
        const QD_G char *pr = (const QD_G char *)(gw.wpot + (long)c0 * gw.M);
        vh[t][h] = *(const QD_G double *)(pr + vb);
      }
  };
#pragma unroll
  for (int d = 0; d < DEPTH - 1; d++) load(d, abuf[d], tbuf[d]);
  stage();
  for (int t0 = 0; t0 < ntiles; t0 += DEPTH) {
#pragma unroll
    for (int d = 0; d < DEPTH; d++) {
      const int tile = t0 + d;
      double vh[NT][2];
      if (OBS && tile < ntiles && tile % nch == nch - 1) load_pot(tile, vh);
      load(tile + DEPTH - 1, abuf[(d + DEPTH - 1) % DEPTH], tbuf[(d + DEPTH - 1) % DEPTH]);
      if (tile < ntiles) compute(tile, abuf[d], tbuf[d], vh);
    }
  }
}

typedef double v2d __attribute__((ext_vector_type(2)));
__device__ __forceinline__ v2d buffer_load_v2d(__amdgpu_buffer_rsrc_t r, uint32_t voff) {
  return __builtin_bit_cast(v2d, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, 0, 0));
}


// FP64 accumulators of one paired-row strip: re and im of 4 NT outputs per
// lane; item it of tile t holds column acc_col<MIXED>(t, lk, it).
template <int NT> struct PairAcc {
  v4d r[NT], i[NT];
  __device__ __forceinline__ void reset() {
#pragma unroll
    for (int t = 0; t < NT; t++) r[t] = i[t] = v4d{0.0, 0.0, 0.0, 0.0};
  }
  __device__ __forceinline__ double re(int t, int it) const { return r[t][it]; }
  __device__ __forceinline__ double im(int t, int it) const { return i[t][it]; }
};

// Multiply-accumulate of one k-chunk (kCh = 4 k-steps, 16 k) of a paired-row
// strip: the re rows (ar) and im rows (ai) of the lane's point for k-steps
// c kCh + q, against the B fragment of the same k-steps.
//   FP64: one v_mfma_f64_16x16x4f64 each for re and im per column tile and
//         k-step.
//   MIXED: a = ah + al and B = bh + bl split into FP32 hi and lo parts.
//     hi*hi: one v_mfma_f32_16x16x4f32 each for re and im per tile and
//            k-step, B hi from gw.Bh, summed in FP32 over the chunk.
//     hi*lo + lo*hi: one v_mfma_f32_16x16x32_bf16 each for re and im per
//            tile and chunk.  Lane (lr, lk) holds the 4 k of its k-rows
//            (k = 16 c + lk + 4 q) and the MFMA pairs its 8 operand slots
//            with the same slots of B: slots 0..3 ah(k_q) against bl(k_q),
//            slots 4..7 al(k_q) against bh(k_q) (the staged table gw.Bx holds
//            those B values as bfloat16, 16 bytes per lane and column).  The
//            terms are 2^-24 of the product, so bfloat16's 8 bits leave an
//            error of 2^-32 of it, below the FP32 rounding of hi*hi.
//     Both FP32 sums of the chunk are added into the FP64 accumulators.
template <int NT, int kCh, bool MIXED>
__device__ __forceinline__ void paired_mac(const GemmWave &gw, int c, int lk, int lr, const double (&ar)[kCh],
                                           const double (&ai)[kCh], PairAcc<NT> &acc) {
  if constexpr (!MIXED) {
#pragma unroll
    for (int q = 0; q < kCh; q++) {
      if (c * kCh + q < gw.ksteps) {
        const double *brow = gw.Bs + (4 * (c * kCh + q) + lk) * gw.Ws + lr;
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const double bv = brow[t * 16];
          acc.r[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, ar[q], acc.r[t], 0, 0, 0);
          acc.i[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, ai[q], acc.i[t], 0, 0, 0);
        }
      }
    }
  } else {
    static_assert(kCh == 4, "the compensation MFMA takes 4 k-steps per lane");
    float hrb[4], lrb[4], hib[4], lib[4];
    v4f hr[NT], hi[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) hr[t] = hi[t] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < kCh; q++) {
      const float arh = (float)ar[q], aih = (float)ai[q];
      hrb[q] = arh;
      lrb[q] = (float)(ar[q] - (double)arh);
      hib[q] = aih;
      lib[q] = (float)(ai[q] - (double)aih);
      if (c * kCh + q < gw.ksteps) {
        const float *brow = gw.Bh + (4 * (c * kCh + q) + lk) * gw.Ws + lr;
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const float bh = brow[t * 16];
          hr[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(bh, arh, hr[t], 0, 0, 0);
          hi[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(bh, aih, hi[t], 0, 0, 0);
        }
      }
    }
    // operand rows past K load as zero, and B rows past kp are staged as zero
    const v8bf xa_r = bf16_operand(hrb, lrb);
    const v8bf xa_i = bf16_operand(hib, lib);
    const v4u *xrow = gw.Bx + (4 * c + lk) * (NT * 16) + lr;
    const v4f zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const v8bf xb = __builtin_bit_cast(v8bf, xrow[t * 16]);
      const v4f xr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xb, xa_r, zero, 0, 0, 0);
      const v4f xi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xb, xa_i, zero, 0, 0, 0);
#pragma unroll
      for (int it = 0; it < 4; it++) {
        acc.r[t][it] += (double)hr[t][it] + (double)xr[it];
        acc.i[t][it] += (double)hi[t][it] + (double)xi[it];
      }
    }
  }
}

// Forward Legendre product with paired rows (FP64): lane (lr, lk) loads the
// complex point p0 + lr of k-step row 4 c + lk as one 16-byte value (its re
// and im rows), and its phase the same way, so the phase multiply stays in
// the lane.  The re rows of a strip of 16 points feed one MFMA and the im
// rows a second, both with the same B fragment from LDS; the two
// accumulators hold re and im of the same outputs, stored as 16 bytes.
// Loads and masking as in gemm_wave (buffer range checks zero rows k >= K).
// MIXED: the products in FP32 hi/lo (paired_mac), everything else as FP64.
template <int NT, int kCh, int DEPTH, bool MIXED, typename Stage>
__device__ __forceinline__ void gemm_wave_fwd_paired(const GemmWave &gw, long strip0, int nstrips, int sstride,
                                                     Stage &&stage) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const long npts = gw.rows / 2;
  const int nch = (gw.ksteps + kCh - 1) / kCh;
  int ns = 0;
  for (int s = 0; s < nstrips; s++)
    if ((strip0 + (long)s * sstride) * 16 < npts) ns = s + 1;
  const int ntiles = ns * nch;
  if (ntiles <= 0) {
    stage();
    return;
  }
  v2d abuf[DEPTH][kCh], tbuf[DEPTH][kCh];
  auto pt_of = [&](int tile) { return (strip0 + (long)(tile / nch) * sstride) * 16 + lr; };
  const uint32_t abytes = (uint32_t)((unsigned long)gw.K * (unsigned long)gw.lda * 8ul);
  const __amdgpu_buffer_rsrc_t rA = make_rsrc(gw.A, abytes);
  const __amdgpu_buffer_rsrc_t rT = make_rsrc((const void *)gw.vph, abytes);
  const uint32_t kstride = (uint32_t)gw.lda * 32u;   // bytes between k and k + 4
  auto load = [&](int tile, v2d(&a)[kCh], v2d(&tp)[kCh]) {
    tile = min(tile, ntiles - 1);
    const int c = tile % nch;
    const long p = pt_of(tile);
    const long pc = p < npts ? p : npts - 1;   // points past the end reread the last one
    const uint32_t v0 = ((uint32_t)(4 * c * kCh + lk) * (uint32_t)gw.lda + 2u * (uint32_t)pc) * 8u;
#pragma unroll
    for (int q = 0; q < kCh; q++) {
      a[q] = buffer_load_v2d(rA, v0 + q * kstride);
      tp[q] = buffer_load_v2d(rT, v0 + q * kstride);
    }
  };
  PairAcc<NT> acc;
  acc.reset();
  auto compute = [&](int tile, const v2d(&an)[kCh], const v2d(&tn)[kCh]) {
    const int c = tile % nch;
    const long p = pt_of(tile);
    if (c == 0) acc.reset();
    double ar[kCh], ai[kCh];
#pragma unroll
    for (int q = 0; q < kCh; q++) {
      const v2d av = an[q], tv = tn[q];
      ar[q] = fma(-av.y, tv.y, av.x * tv.x);
      ai[q] = fma(av.x, tv.y, av.y * tv.x);
    }
    paired_mac<NT, kCh, MIXED>(gw, c, lk, lr, ar, ai, acc);
    if (c == nch - 1) {   // epilogue of the strip: streaming 16-byte stores
      const bool pok = p < npts;
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int it = 0; it < 4; it++) {
          const int col = acc_col<MIXED>(t, lk, it);
          if (pok && col < gw.N)
            __builtin_nontemporal_store(v2d{acc.re(t, it), acc.im(t, it)}, (v2d *)(gw.C + (long)col * gw.ldc + 2 * p));
        }
    }
  };
#pragma unroll
  for (int d = 0; d < DEPTH - 1; d++) load(d, abuf[d], tbuf[d]);
  stage();
  for (int t0 = 0; t0 < ntiles; t0 += DEPTH) {
#pragma unroll
    for (int d = 0; d < DEPTH; d++) {
      const int tile = t0 + d;
      load(tile + DEPTH - 1, abuf[(d + DEPTH - 1) % DEPTH], tbuf[(d + DEPTH - 1) % DEPTH]);
      if (tile < ntiles) compute(tile, abuf[d], tbuf[d]);
    }
  }
}

// Backward Legendre product with paired rows (FP64): as gemm_wave_fwd_paired
// without the phase, plus the potential-energy epilogue: each lane owns whole
// points, so it loads the w_k V values of its point (4 NT per strip, issued
// before the prefetch) and adds w V (re^2 + im^2) of every output into ob.
// MIXED: the products in FP32 hi/lo (paired_mac).
template <int NT, int kCh, int DEPTH, bool MIXED, typename Stage>
__device__ __forceinline__ void gemm_wave_bwd_paired(const GemmWave &gw, long strip0, int nstrips, int sstride,
                                                     double &ob, Stage &&stage) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const long npts = gw.rows / 2;
  const int nch = (gw.ksteps + kCh - 1) / kCh;
  int ns = 0;
  for (int s = 0; s < nstrips; s++)
    if ((strip0 + (long)s * sstride) * 16 < npts) ns = s + 1;
  const int ntiles = ns * nch;
  if (ntiles <= 0) {
    stage();
    return;
  }
  v2d abuf[DEPTH][kCh];
  auto pt_of = [&](int tile) { return (strip0 + (long)(tile / nch) * sstride) * 16 + lr; };
  // operand rows k < K only: rows K.. of the padded product read as zero
  const uint32_t abytes = (uint32_t)((unsigned long)gw.K * (unsigned long)gw.lda * 8ul);
  const __amdgpu_buffer_rsrc_t rA = make_rsrc(gw.A, abytes);
  const uint32_t kstride = (uint32_t)gw.lda * 32u;
  auto load = [&](int tile, v2d(&a)[kCh]) {
    tile = min(tile, ntiles - 1);
    const int c = tile % nch;
    const long p = pt_of(tile);
    const long pc = p < npts ? p : npts - 1;
    const uint32_t v0 = ((uint32_t)(4 * c * kCh + lk) * (uint32_t)gw.lda + 2u * (uint32_t)pc) * 8u;
#pragma unroll
    for (int q = 0; q < kCh; q++) a[q] = buffer_load_v2d(rA, v0 + q * kstride);
  };
  // w_k V of this lane's point for its 4 NT output columns (the table holds
  // NT * 16 theta rows, zero past N: no clamp on the column)
  auto load_pot = [&](int tile, double(&vh)[NT][4]) {
    const long p = pt_of(tile);
    const long pc = p < npts ? p : npts - 1;
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int it = 0; it < 4; it++) vh[t][it] = gw.wpot[(long)acc_col<MIXED>(t, lk, it) * gw.M + pc];
  };
  PairAcc<NT> acc;
  acc.reset();
  auto compute = [&](int tile, const v2d(&an)[kCh], const double(&vh)[NT][4]) {
    const int c = tile % nch;
    const long p = pt_of(tile);
    if (c == 0) acc.reset();
    double ar[kCh], ai[kCh];
#pragma unroll
    for (int q = 0; q < kCh; q++) {
      ar[q] = an[q].x;
      ai[q] = an[q].y;
    }
    paired_mac<NT, kCh, MIXED>(gw, c, lk, lr, ar, ai, acc);
    if (c == nch - 1) {
      const bool pok = p < npts;
      // the mixed form loads the potentials here, after its FP32 partial
      // sums are gone (issued ahead of the chunk they would not fit beside them)
      double vm[NT][4];
      if constexpr (MIXED) load_pot(tile, vm);
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int it = 0; it < 4; it++) {
          const int col = acc_col<MIXED>(t, lk, it);
          const double re = acc.re(t, it), im = acc.im(t, it);
          if (pok && col < gw.N)
            __builtin_nontemporal_store(v2d{re, im}, (v2d *)(gw.C + (long)col * gw.ldc + 2 * p));
          const double wv = MIXED ? vm[t][it] : vh[t][it];
          ob += (pok ? re * re + im * im : 0.0) * wv;   // w V = 0 for col >= N
        }
    }
  };
#pragma unroll
  for (int d = 0; d < DEPTH - 1; d++) load(d, abuf[d]);
  stage();
  for (int t0 = 0; t0 < ntiles; t0 += DEPTH) {
#pragma unroll
    for (int d = 0; d < DEPTH; d++) {
      const int tile = t0 + d;
      double vh[NT][4];
      if (!MIXED && tile < ntiles && tile % nch == nch - 1) load_pot(tile, vh);
      load(tile + DEPTH - 1, abuf[(d + DEPTH - 1) % DEPTH]);
      if (tile < ntiles) compute(tile, abuf[d], vh);
    }
  }
}

}  // namespace qd
