// kernels_leg_split.hip - the fused Legendre step (kernels_leg_step.hip) with
// both products split by parity, half the matrix-core work.
//
// On a Gauss-Legendre grid the nodes come in pairs x_k' = -x_k, k' = nt - 1 - k,
// with equal weights, and the normalised associated Legendre functions have
// parity P_l^omega(-x) = (-1)^(l - omega) P_l^omega(x).  With nh = nt / 2:
//   backward (omegawavepacket.cu:221-240):
//     E_k = sum_{l - omega even} P_l(x_k) FBR_l,  O_k = sum_{l - omega odd} P_l(x_k) FBR_l  (k < nh)
//     psi_k = E_k + O_k,  psi_k' = E_k - O_k
//   forward (omegawavepacket.cu:200-219), phi = d1 d2 exp(-i dt V) psi:
//     FBR'_l = sum_{k < nh} w_k P_l(x_k) (phi_k + phi_k')   (l - omega even)
//     FBR'_l = sum_{k < nh} w_k P_l(x_k) (phi_k - phi_k')   (l - omega odd)
// Each of the four products has half the terms of a direct one, so the pair of
// transforms costs half the MFMAs.  In the paired-row layout E_k and O_k of a
// point sit in the same lane, item it of tile t holding k = 16 t + lk + 4 it,
// which is also the lane's forward operand of k-step 4 t + it: everything
// between the products is lane-local.  The job set-up checks the caller's
// nodes, weights and tables for this symmetry (DeviceJob) and uses the
// direct form otherwise.  The results agree with the direct sums to rounding.
//
// Mixed form (MIXED): every product as FP32 hi*hi MFMAs plus one bfloat16
// compensation MFMA per 16 k (paired_mac), flushed into FP64 accumulators.
// The FP32 16x16x4 accumulator holds column 16 t + 4 lk + it instead of
// 16 t + lk + 4 it, so a lane's backward outputs are theta rows
// acc_col<true>; the forward tables are stored with their rows in the order
// the forward MFMAs consume them (row 16 c + 4 q + g holds node
// 16 c + 4 g + q), so the hand-off stays in the lane.
//
// At 3 theta tiles per half (96 theta points, FP64) the forward products run
// in two sweeps over the theta chunks, even l then odd l, so that one set of
// forward accumulators is live at a time; the even sweep leaves the odd
// product's operands in LDS (split_wave).
//
// Modes, as the direct fused kernel: fused (backward product of the pending
// step, potential energy, flux stencil rows, phase, forward product), FIN
// (backward product alone, the DVR state stored, finishing a step sequence),
// and forward-only for jobs whose state is in DVR.
#include "gemm_wave.h"

namespace qd {

// One k-chunk (4 k-steps, 16 k) of a mixed product from B fragments already
// read from LDS: FP32 hi parts bh[q][t0 + t] of k-step q and the bfloat16
// compensation entries bx[t0 + t] (paired_mac's arithmetic: hi*hi summed in
// FP32 over the chunk, hi*lo + lo*hi on one bf16 MFMA, both added into the
// FP64 accumulators).
template <int NT, int NB>
__device__ __forceinline__ void split_mac_mixed(const float (&bh)[4][NB], const v4u (&bx)[NB], int t0,
                                                const double (&ar)[4], const double (&ai)[4], PairAcc<NT> &acc) {
  float hrb[4], lrb[4], hib[4], lib[4];
  v4f hr[NT], hi[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) hr[t] = hi[t] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const float arh = (float)ar[q], aih = (float)ai[q];
    hrb[q] = arh;
    lrb[q] = (float)(ar[q] - (double)arh);
    hib[q] = aih;
    lib[q] = (float)(ai[q] - (double)aih);
#pragma unroll
    for (int t = 0; t < NT; t++) {
      hr[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(bh[q][t0 + t], arh, hr[t], 0, 0, 0);
      hi[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(bh[q][t0 + t], aih, hi[t], 0, 0, 0);
    }
  }
  const v8bf xa_r = bf16_operand(hrb, lrb);
  const v8bf xa_i = bf16_operand(hib, lib);
  const v4f zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const v8bf xb = __builtin_bit_cast(v8bf, bx[t0 + t]);
    const v4f xr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xb, xa_r, zero, 0, 0, 0);
    const v4f xi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xb, xa_i, zero, 0, 0, 0);
#pragma unroll
    for (int it = 0; it < 4; it++) {
      acc.r[t][it] += (double)hr[t][it] + (double)xr[it];
      acc.i[t][it] += (double)hi[t][it] + (double)xi[it];
    }
  }
}

// Store of a forward product's output (FBR coefficients, read once by the
// next pass): streaming when the job's arrays are larger than the Infinity
// Cache, plain otherwise so that the next pass finds them there.
__device__ __forceinline__ void fbr_store(int stream, v2d val, v2d *p) {
  if (stream) __builtin_nontemporal_store(val, p);
  else *p = val;
}

// Device view of one block of a split job.
struct SplitView {
  // mixed form: FP32 hi tables and bfloat16 compensation tables in LDS
  // (same rows and columns as the FP64 ones; compensation rows of 16 k)
  const float *BE1h, *BO1h, *BE2h, *BO2h;
  const v4u *BE1x, *BO1x, *BE2x, *BO2x;
  const double *fbr;        // FBR slab l = omega_b of block b, doubles, rows of 2M
  int nl;                   // l_max - omega_b + 1 (operand rows; rows past it load as zero)
  QD_G double2 *psi;        // DVR block b, [nt][M]
  const QD_G double *vph;   // phase table as doubles, rows of 2M
  const QD_G double *wpot;  // w_k V rows [bwd_np][M]
  int nt, nh, M, n1, j0, j1;// surface rows [j0, j1) (fused), all rows (FIN)
  int stream;               // FBR output by streaming stores (arrays larger than the Infinity Cache)
  int ne, no;               // even / odd l columns of this block
  const double *BE1, *BO1;  // LDS: P_E [16 TE][WsH], P_O [16 TO][WsH]  (backward, rows m, columns k)
  v2d *odd;                 // LDS, two-sweep form: this wave's odd-sweep operands [TH][4][64 lanes]
  const double *BE2, *BO2;  // LDS: wP_E [16 TH][WsE], wP_O [16 TH][WsO] (forward, rows k, columns m)
};

// One wave's strips in the fused (FIN = false) or backward-only (FIN) mode,
// one wave per SIMD.  A strip's tiles: TE k-chunks of the even backward
// product, TO of the odd one (FBR rows l = omega + 2 m and omega + 2 m + 1),
// then TH chunks that form psi_k, psi_k' from E and O (potential energy,
// stored rows) and, fused, feed both forward products.  With 3 or more
// theta tiles per half (TWO) the forward products take two sweeps over the
// theta chunks, the even one with the potential energy and stored rows,
// then the odd one, which forms psi and phi again from E and O: one set of
// forward accumulators is live at a time, at the cost of reading the phase
// rows twice.  Padded with idle tiles to a multiple of lcm(D, 2), unrolled
// and fenced tile by tile; each fence region issues tile tau's MFMAs with
// tile tau + 1's preparation (B fragments from LDS, operands from the
// ring), as step_wave_paired does.
template <int TH, int TE, int TO, int D, bool FIN, bool MIXED, typename Stage>
__device__ __forceinline__ void split_wave(const SplitView &v, long strip0, int nstrips, int sstride, double &ob,
                                           Stage &&stage) {
  constexpr int kCh = 4;
  constexpr bool TWO = TH >= 3;
  constexpr int TB = TE + TO;                                // backward tiles
  constexpr int T0 = TB + TH + (TWO && !FIN ? TH : 0);       // tiles of a strip
  constexpr int LCM = D % 2 == 0 ? D : 2 * D;
  constexpr int T = (T0 + LCM - 1) / LCM * LCM;
  constexpr int WsH = lds_stride(TH), WsE = lds_stride(TE), WsO = lds_stride(TO);
  constexpr int NF = TWO ? (TE > TO ? TE : TO) : TE + TO;   // B fragment tiles of a forward tile
  constexpr int NB = NF > TH ? NF : TH;
  constexpr int OO = TWO ? 0 : TE;                           // first fragment tile of the odd product
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const long npts = v.M;
  const uint32_t lda = 2u * (uint32_t)v.M;   // doubles per row of the FBR, phase and psi arrays
  int ns = 0;
  for (int s = 0; s < nstrips; s++)
    if ((strip0 + (long)s * sstride) * 16 < npts) ns = s + 1;
  if (ns <= 0) {
    stage();
    return;
  }
  struct Slot {
    v2d x[kCh], y[kCh];   // backward: FBR values (x); forward: phase at k (x) and k' (y)
    double w[2 * kCh];    // w V at k (0..3) and k' (4..7)
  };
  struct Op {
    double ar[kCh], ai[kCh], br[kCh], bi[kCh];   // backward: FBR (a); forward: even (a) and odd (b) operands
    double bv[kCh][NB];   // FP64 B fragments
    float bh[kCh][NB];    // mixed: FP32 hi fragments and compensation entries
    v4u bx[NB];
  };
  Slot ring[D];
  Op op[2];
  const __amdgpu_buffer_rsrc_t rA = make_rsrc(v.fbr, (uint32_t)((unsigned long)v.nl * lda * 8ul));
  const __amdgpu_buffer_rsrc_t rT = make_rsrc((const void *)v.vph, (uint32_t)((unsigned long)v.nt * lda * 8ul));
  // theta row of item q of a lane's backward tile c (the operand of forward
  // k-step 4 c + q) and its partner
  auto kk = [&](int c, int q) { return acc_col<MIXED>(c, lk, q); };
  auto kp = [&](int c, int q) { return 2 * v.nh - 1 - kk(c, q); };
  // forward tile tau: second sweep (the odd product, TWO) and theta chunk
  auto second = [&](int tau) { return TWO && tau >= TB + TH; };
  auto chunk = [&](int tau) { return tau - TB - (TWO && tau >= TB + TH ? TH : 0); };
  auto load = [&](int tau, long st, Slot &r) {
    const long p = st * 16 + lr;
    const uint32_t pc = (uint32_t)(p < npts ? p : npts - 1);   // points past the end reread the last one
    if (tau < TB) {
      const int c = tau < TE ? tau : tau - TE, par = tau < TE ? 0 : 1;
#pragma unroll
      for (int q = 0; q < kCh; q++) {
        const uint32_t row = 2u * (uint32_t)(16 * c + 4 * q + lk) + (uint32_t)par;   // l - omega
        r.x[q] = buffer_load_v2d(rA, (row * lda + 2u * pc) * 8u);
      }
    } else if (tau < T0) {
      const int c = chunk(tau);
#pragma unroll
      for (int q = 0; q < kCh; q++) {
        // partner rows below 0 (k past nh on small grids) wrap to huge
        // offsets: the range check returns zero (their psi is zero anyway)
        if (!FIN && !second(tau)) {
          r.x[q] = buffer_load_v2d(rT, ((uint32_t)kk(c, q) * lda + 2u * pc) * 8u);
          r.y[q] = buffer_load_v2d(rT, ((uint32_t)kp(c, q) * lda + 2u * pc) * 8u);
        }
        if (!second(tau)) {
          const int k1 = kk(c, q), k2 = max(kp(c, q), 0);
          r.w[q] = v.wpot[(long)k1 * v.M + pc];
          r.w[kCh + q] = v.wpot[(long)k2 * v.M + pc];
        }
      }
    }
  };
  PairAcc<TH> aE, aO;   // E_k, O_k: item it of tile t at k = 16 t + lk + 4 it
  PairAcc<TE> bE;       // forward outputs, l = omega + 2 m, m = 16 t + lk + 4 it
  PairAcc<TO> bO;       // l = omega + 2 m + 1
  // B fragments of tile tau from LDS: read one tile ahead by prep, or (TWO:
  // fewer registers) in the tile's own fence region by mma
  auto frag = [&](int tau, Op &o) {
    if (tau < TB) {
      const int c = tau < TE ? tau : tau - TE;
#pragma unroll
      for (int q = 0; q < kCh; q++) {
        if (!MIXED) {
          const double *B = tau < TE ? v.BE1 : v.BO1;
#pragma unroll
          for (int t = 0; t < TH; t++) o.bv[q][t] = B[(16 * c + 4 * q + lk) * WsH + lr + 16 * t];
        } else {
          const float *Bh = tau < TE ? v.BE1h : v.BO1h;
#pragma unroll
          for (int t = 0; t < TH; t++) o.bh[q][t] = Bh[(16 * c + 4 * q + lk) * WsH + lr + 16 * t];
        }
      }
      if (MIXED) {
        const v4u *Bx = tau < TE ? v.BE1x : v.BO1x;
#pragma unroll
        for (int t = 0; t < TH; t++) o.bx[t] = Bx[(4 * c + lk) * (16 * TH) + lr + 16 * t];
      }
    } else if (!FIN && tau < T0) {
      const bool s1 = second(tau), even = !s1, odd = !TWO || s1;
      const int c = chunk(tau);
#pragma unroll
      for (int q = 0; q < kCh; q++) {
        if (!MIXED) {
          if (even) {
#pragma unroll
            for (int t = 0; t < TE; t++) o.bv[q][t] = v.BE2[(16 * c + 4 * q + lk) * WsE + lr + 16 * t];
          }
          if (odd) {
#pragma unroll
            for (int t = 0; t < TO; t++) o.bv[q][OO + t] = v.BO2[(16 * c + 4 * q + lk) * WsO + lr + 16 * t];
          }
        } else {   // table row 16 c + 4 q + lk holds node kk(c, q)
          if (even) {
#pragma unroll
            for (int t = 0; t < TE; t++) o.bh[q][t] = v.BE2h[(16 * c + 4 * q + lk) * WsE + lr + 16 * t];
          }
          if (odd) {
#pragma unroll
            for (int t = 0; t < TO; t++) o.bh[q][OO + t] = v.BO2h[(16 * c + 4 * q + lk) * WsO + lr + 16 * t];
          }
        }
      }
      if (MIXED) {
        if (even) {
#pragma unroll
          for (int t = 0; t < TE; t++) o.bx[t] = v.BE2x[(4 * c + lk) * (16 * TE) + lr + 16 * t];
        }
        if (odd) {
#pragma unroll
          for (int t = 0; t < TO; t++) o.bx[OO + t] = v.BO2x[(4 * c + lk) * (16 * TO) + lr + 16 * t];
        }
      }
    }
  };
  auto prep = [&](int tau, long st, bool surf, const Slot &r, Op &o) {
    const long p = st * 16 + lr;
    const bool pok = p < npts;
    if (tau < TB) {
#pragma unroll
      for (int q = 0; q < kCh; q++) {
        o.ar[q] = r.x[q].x;
        o.ai[q] = r.x[q].y;
      }
    } else if (tau < T0 && second(tau)) {
      // odd sweep: the operands phi_k - phi_k' that the even sweep left in LDS
      const int c = chunk(tau);
#pragma unroll
      for (int q = 0; q < kCh; q++) {
        const v2d d = v.odd[(c * kCh + q) * 64 + lane];
        o.ar[q] = d.x;
        o.ai[q] = d.y;
      }
    } else if (tau < T0) {
      const int c = chunk(tau);
      const v4d er = aE.r[c], ei = aE.i[c], orr = aO.r[c], oi = aO.i[c];
#pragma unroll
      for (int q = 0; q < kCh; q++) {
        const double pr = er[q] + orr[q], pi = ei[q] + oi[q];    // psi_k
        const double qr = er[q] - orr[q], qi = ei[q] - oi[q];    // psi_k'
        // w V = 0 past nt; psi = 0 at k >= nh (P columns zero there)
        ob += (pok ? pr * pr + pi * pi : 0.0) * r.w[q];
        ob += (pok ? qr * qr + qi * qi : 0.0) * r.w[kCh + q];
        if (surf && kk(c, q) < v.nh) {
          *(QD_G v2d *)(v.psi + (long)kk(c, q) * v.M + p) = v2d{pr, pi};
          *(QD_G v2d *)(v.psi + (long)kp(c, q) * v.M + p) = v2d{qr, qi};
        }
        if (!FIN) {
          const v2d t1 = r.x[q], t2 = r.y[q];
          const double fr = fma(-pi, t1.y, pr * t1.x), fi = fma(pr, t1.y, pi * t1.x);   // phi_k
          const double gr = fma(-qi, t2.y, qr * t2.x), gi = fma(qr, t2.y, qi * t2.x);   // phi_k'
          o.ar[q] = fr + gr;
          o.ai[q] = fi + gi;
          if (TWO) {   // the odd sweep's operands, kept for it in LDS
            v.odd[(c * kCh + q) * 64 + lane] = v2d{fr - gr, fi - gi};
          } else {
            o.br[q] = fr - gr;
            o.bi[q] = fi - gi;
          }
        }
      }
    }
    if (!TWO) frag(tau, o);
  };
  auto mma = [&](int tau, long st, Op &o) {
    if (TWO) frag(tau, o);
    if (tau < TB) {
      PairAcc<TH> &acc = tau < TE ? aE : aO;
      if (tau == 0) {
        aE.reset();
        aO.reset();
      }
      if constexpr (MIXED) {
        split_mac_mixed<TH, NB>(o.bh, o.bx, 0, o.ar, o.ai, acc);
      } else {
#pragma unroll
        for (int q = 0; q < kCh; q++)
#pragma unroll
          for (int t = 0; t < TH; t++) {
            acc.r[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(o.bv[q][t], o.ar[q], acc.r[t], 0, 0, 0);
            acc.i[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(o.bv[q][t], o.ai[q], acc.i[t], 0, 0, 0);
          }
      }
    } else if (!FIN && tau < T0) {
      const bool s1 = second(tau), even = !s1, odd = !TWO || s1;
      const int c = chunk(tau);
      if (c == 0) {
        if (even) bE.reset();
        if (odd) bO.reset();
      }
      if constexpr (MIXED) {
        if (even) split_mac_mixed<TE, NB>(o.bh, o.bx, 0, o.ar, o.ai, bE);
        if (odd) split_mac_mixed<TO, NB>(o.bh, o.bx, OO, TWO ? o.ar : o.br, TWO ? o.ai : o.bi, bO);
      } else {
#pragma unroll
        for (int q = 0; q < kCh; q++) {
          if (even) {
#pragma unroll
            for (int t = 0; t < TE; t++) {
              bE.r[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(o.bv[q][t], o.ar[q], bE.r[t], 0, 0, 0);
              bE.i[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(o.bv[q][t], o.ai[q], bE.i[t], 0, 0, 0);
            }
          }
          if (odd) {
#pragma unroll
            for (int t = 0; t < TO; t++) {
              const double xr = TWO ? o.ar[q] : o.br[q], xi = TWO ? o.ai[q] : o.bi[q];
              bO.r[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(o.bv[q][OO + t], xr, bO.r[t], 0, 0, 0);
              bO.i[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(o.bv[q][OO + t], xi, bO.i[t], 0, 0, 0);
            }
          }
        }
      }
      if (c == TH - 1) {   // streaming 16-byte stores of the sweep's parities
        const long p = st * 16 + lr;
        const bool pok = p < npts;
        if (even) {
#pragma unroll
          for (int t = 0; t < TE; t++)
#pragma unroll
            for (int it = 0; it < 4; it++) {
              const int m = acc_col<MIXED>(t, lk, it);
              if (pok && m < v.ne)
                fbr_store(v.stream, v2d{bE.re(t, it), bE.im(t, it)},
                                            (v2d *)(v.fbr + (long)(2 * m) * lda + 2 * p));
            }
        }
        if (odd) {
#pragma unroll
          for (int t = 0; t < TO; t++)
#pragma unroll
            for (int it = 0; it < 4; it++) {
              const int m = acc_col<MIXED>(t, lk, it);
              if (pok && m < v.no)
                fbr_store(v.stream, v2d{bO.re(t, it), bO.im(t, it)},
                                            (v2d *)(v.fbr + (long)(2 * m + 1) * lda + 2 * p));
            }
        }
      }
    }
  };
  auto on_surface = [&](long st) {   // this lane's point lies on a stored row (r row p / n1)
    const int pj = (int)(st * 16 + lr) / v.n1;
    return st * 16 + lr < npts && pj >= v.j0 && pj < v.j1;
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

// Forward product alone (a job whose state is in DVR): per strip the TH
// chunks load psi and the phase at k and k', form the even and odd operands
// as the fused mode does, and feed both forward products; arithmetic
// identical to the fused mode's forward half.
template <int TH, int TE, int TO, bool MIXED>
__device__ __forceinline__ void split_forward_wave(const SplitView &v, long strip0, int nstrips, int sstride) {
  constexpr int kCh = 4;
  constexpr int WsE = lds_stride(TE), WsO = lds_stride(TO);
  constexpr bool TWO = TH >= 3;   // even and odd products in separate sweeps, as the fused mode
  constexpr int NB = TE + TO;
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const long npts = v.M;
  const uint32_t lda = 2u * (uint32_t)v.M;
  const __amdgpu_buffer_rsrc_t rP =
      make_rsrc((const void *)v.psi, (uint32_t)((unsigned long)v.nt * lda * 8ul));
  const __amdgpu_buffer_rsrc_t rT = make_rsrc((const void *)v.vph, (uint32_t)((unsigned long)v.nt * lda * 8ul));
  for (int s = 0; s < nstrips; s++) {
    const long st = strip0 + (long)s * sstride;
    if (st * 16 >= npts) break;
    const long p = st * 16 + lr;
    const bool pok = p < npts;
    const uint32_t pc = (uint32_t)(pok ? p : npts - 1);
#pragma unroll
    for (int sw = 0; sw < (TWO ? 2 : 1); sw++) {
      const bool even = !TWO || sw == 0, odd = !TWO || sw == 1;
      PairAcc<TE> bE;
      PairAcc<TO> bO;
      bE.reset();
      bO.reset();
#pragma unroll
      for (int c = 0; c < TH; c++) {
        v2d x1[kCh], x2[kCh], t1[kCh], t2[kCh];
#pragma unroll
        for (int q = 0; q < kCh; q++) {
          const uint32_t k1 = (uint32_t)acc_col<MIXED>(c, lk, q), k2 = (uint32_t)(2 * v.nh - 1) - k1;
          x1[q] = buffer_load_v2d(rP, (k1 * lda + 2u * pc) * 8u);
          x2[q] = buffer_load_v2d(rP, (k2 * lda + 2u * pc) * 8u);
          t1[q] = buffer_load_v2d(rT, (k1 * lda + 2u * pc) * 8u);
          t2[q] = buffer_load_v2d(rT, (k2 * lda + 2u * pc) * 8u);
        }
        double ar[kCh], ai[kCh], br[kCh], bi[kCh];
#pragma unroll
        for (int q = 0; q < kCh; q++) {
          // rows k >= nh are zero in the fused mode (E and O vanish there)
          const bool kv = acc_col<MIXED>(c, lk, q) < v.nh;
          const double pr = kv ? x1[q].x : 0.0, pi = kv ? x1[q].y : 0.0;
          const double qr = kv ? x2[q].x : 0.0, qi = kv ? x2[q].y : 0.0;
          const double fr = fma(-pi, t1[q].y, pr * t1[q].x), fi = fma(pr, t1[q].y, pi * t1[q].x);
          const double gr = fma(-qi, t2[q].y, qr * t2[q].x), gi = fma(qr, t2[q].y, qi * t2[q].x);
          ar[q] = fr + gr;
          ai[q] = fi + gi;
          br[q] = fr - gr;
          bi[q] = fi - gi;
        }
        if constexpr (MIXED) {
          float bh[kCh][NB];
          v4u bx[NB];
#pragma unroll
          for (int q = 0; q < kCh; q++) {
#pragma unroll
            for (int t = 0; t < TE; t++) bh[q][t] = v.BE2h[(16 * c + 4 * q + lk) * WsE + lr + 16 * t];
#pragma unroll
            for (int t = 0; t < TO; t++) bh[q][TE + t] = v.BO2h[(16 * c + 4 * q + lk) * WsO + lr + 16 * t];
          }
#pragma unroll
          for (int t = 0; t < TE; t++) bx[t] = v.BE2x[(4 * c + lk) * (16 * TE) + lr + 16 * t];
#pragma unroll
          for (int t = 0; t < TO; t++) bx[TE + t] = v.BO2x[(4 * c + lk) * (16 * TO) + lr + 16 * t];
          if (even) split_mac_mixed<TE, NB>(bh, bx, 0, ar, ai, bE);
          if (odd) split_mac_mixed<TO, NB>(bh, bx, TE, br, bi, bO);
        } else {
#pragma unroll
          for (int q = 0; q < kCh; q++) {
            if (even) {
#pragma unroll
              for (int t = 0; t < TE; t++) {
                const double bv = v.BE2[(16 * c + 4 * q + lk) * WsE + lr + 16 * t];
                bE.r[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, ar[q], bE.r[t], 0, 0, 0);
                bE.i[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, ai[q], bE.i[t], 0, 0, 0);
              }
            }
            if (odd) {
#pragma unroll
              for (int t = 0; t < TO; t++) {
                const double bv = v.BO2[(16 * c + 4 * q + lk) * WsO + lr + 16 * t];
                bO.r[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, br[q], bO.r[t], 0, 0, 0);
                bO.i[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, bi[q], bO.i[t], 0, 0, 0);
              }
            }
          }
        }
      }
      if (even) {
#pragma unroll
        for (int t = 0; t < TE; t++)
#pragma unroll
          for (int it = 0; it < 4; it++) {
            const int m = acc_col<MIXED>(t, lk, it);
            if (pok && m < v.ne)
              fbr_store(v.stream, v2d{bE.re(t, it), bE.im(t, it)},
                                          (v2d *)(v.fbr + (long)(2 * m) * lda + 2 * p));
          }
      }
      if (odd) {
#pragma unroll
        for (int t = 0; t < TO; t++)
#pragma unroll
          for (int it = 0; it < 4; it++) {
            const int m = acc_col<MIXED>(t, lk, it);
            if (pok && m < v.no)
              fbr_store(v.stream, v2d{bO.re(t, it), bO.im(t, it)},
                                          (v2d *)(v.fbr + (long)(2 * m + 1) * lda + 2 * p));
          }
      }
    }
  }
}

// Workgroup (block b, group g): 4 waves, one per SIMD, the strips
// g S 4 + wave + 4 s of block b (the fused kernel's decomposition).  LDS: the
// forward tables wP_E, wP_O, then (fused, FIN) the backward tables P_E, P_O,
// then (two-sweep form) each wave's odd-sweep operands; mixed: the four
// compensation tables (16-byte entries, rows of 16 k), then the four FP32 hi
// tables.
template <int TH, int TE, int TO, bool FIN, bool MIXED>
__global__ __launch_bounds__(4 * 64) void k_leg_split(StepArgs a) {
  constexpr int WAVES = 4, TB = WAVES * 64;
  constexpr int WsH = lds_stride(TH), WsE = lds_stride(TE), WsO = lds_stride(TO);
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
  const bool fused = jb.counters[3] != 0;   // the job's backward transform is pending
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  SplitView v;
  double *BE2 = lds_d;                       // [16 TH][WsE]
  double *BO2 = BE2 + 16 * TH * WsE;         // [16 TH][WsO]
  double *BE1 = BO2 + 16 * TH * WsO;         // [16 TE][WsH]
  double *BO1 = BE1 + 16 * TE * WsH;         // [16 TO][WsH]
  // two-sweep form (3 or more theta tiles per half, FP64): each wave's
  // odd-sweep operands [TH][4][64] after the tables
  v2d *ODD = (v2d *)(BO1 + 16 * TO * WsH) + (long)wave * TH * 4 * 64;
  v4u *BE2x = (v4u *)lds_d;                  // [4 TH][16 TE]
  v4u *BO2x = BE2x + 4 * TH * 16 * TE;       // [4 TH][16 TO]
  v4u *BE1x = BO2x + 4 * TH * 16 * TO;       // [4 TE][16 TH]
  v4u *BO1x = BE1x + 4 * TE * 16 * TH;       // [4 TO][16 TH]
  float *BE2h = (float *)(BO1x + 4 * TO * 16 * TH);   // as the FP64 tables
  float *BO2h = BE2h + 16 * TH * WsE;
  float *BE1h = BO2h + 16 * TH * WsO;
  float *BO1h = BE1h + 16 * TE * WsH;
  const bool need_fwd = !FIN, need_bwd = fused;
  auto stage = [&]() {
    // tables are stored padded to whole 16-row / 16-column chunks
    if constexpr (!MIXED) {
      if (need_fwd) {
        stage_b<TB, 16 * TE>(BE2, jb.sp_wPE + (long)b * 16 * TH * 16 * TE, 16 * TH, 16 * TE, WsE);
        stage_b<TB, 16 * TO>(BO2, jb.sp_wPO + (long)b * 16 * TH * 16 * TO, 16 * TH, 16 * TO, WsO);
      }
      if (need_bwd) {
        stage_b<TB, 16 * TH>(BE1, jb.sp_PE + (long)b * 16 * TE * 16 * TH, 16 * TE, 16 * TH, WsH);
        stage_b<TB, 16 * TH>(BO1, jb.sp_PO + (long)b * 16 * TO * 16 * TH, 16 * TO, 16 * TH, WsH);
      }
    } else {
      if (need_fwd) {
        stage_b<TB, 16 * TE>(BE2h, jb.sp_hwPE + (long)b * 16 * TH * 16 * TE, 16 * TH, 16 * TE, WsE);
        stage_b<TB, 16 * TO>(BO2h, jb.sp_hwPO + (long)b * 16 * TH * 16 * TO, 16 * TH, 16 * TO, WsO);
        stage_b<TB, 16 * TE>(BE2x, (const QD_G v4u *)jb.sp_xwPE + (long)b * 4 * TH * 16 * TE, 4 * TH, 16 * TE,
                             16 * TE);
        stage_b<TB, 16 * TO>(BO2x, (const QD_G v4u *)jb.sp_xwPO + (long)b * 4 * TH * 16 * TO, 4 * TH, 16 * TO,
                             16 * TO);
      }
      if (need_bwd) {
        stage_b<TB, 16 * TH>(BE1h, jb.sp_hPE + (long)b * 16 * TE * 16 * TH, 16 * TE, 16 * TH, WsH);
        stage_b<TB, 16 * TH>(BO1h, jb.sp_hPO + (long)b * 16 * TO * 16 * TH, 16 * TO, 16 * TH, WsH);
        stage_b<TB, 16 * TH>(BE1x, (const QD_G v4u *)jb.sp_xPE + (long)b * 4 * TE * 16 * TH, 4 * TE, 16 * TH,
                             16 * TH);
        stage_b<TB, 16 * TH>(BO1x, (const QD_G v4u *)jb.sp_xPO + (long)b * 4 * TO * 16 * TH, 4 * TO, 16 * TH,
                             16 * TH);
      }
    }
    __syncthreads();
  };
  v.fbr = (const double *)(jb.fbr + ((long)b * jb.L + om) * M);
  v.nl = nl;
  v.psi = jb.psi + b * G;
  v.vph = (const QD_G double *)jb.vphase;
  v.wpot = jb.wpot;
  v.nt = jb.nt;
  v.nh = jb.nt / 2;
  v.M = jb.M;
  v.n1 = jb.n1;
  v.stream = jb.stream_out;
  v.j0 = FIN ? 0 : jb.flux_on ? jb.flux_ns - jb.flux_ng / 2 : 0;
  v.j1 = FIN ? jb.n2 : jb.flux_on ? v.j0 + jb.flux_ng : 0;
  v.ne = (nl + 1) / 2;
  v.no = nl / 2;
  v.BE1 = BE1;
  v.BO1 = BO1;
  v.BE2 = BE2;
  v.BO2 = BO2;
  v.odd = ODD;
  v.BE1h = BE1h;
  v.BO1h = BO1h;
  v.BE2h = BE2h;
  v.BO2h = BO2h;
  v.BE1x = BE1x;
  v.BO1x = BO1x;
  v.BE2x = BE2x;
  v.BO2x = BO2x;
  const int S = jb.step_strips;
  const long strip0 = (long)g * S * WAVES + wave;
  if (!fused) {
    if (!FIN) {
      stage();
      split_forward_wave<TH, TE, TO, MIXED>(v, strip0, S, WAVES);
    }
  } else {
    double ob = 0.0;
    split_wave<TH, TE, TO, (TH >= 3 ? 2 : 3), FIN, MIXED>(v, strip0, S, WAVES, ob, stage);
    ob = wave_sum64(ob);
    if (lane == 0) red[wave] = ob;
    __syncthreads();
    if (threadIdx.x == 0) jb.part_obs[(long)b * jb.step_groups + g] = (red[0] + red[1]) + (red[2] + red[3]);
  }
  // apply the counter advance of the reduction before this kernel
  if (u == 0 && threadIdx.x == 0) advance_counters(jb);
}

// Instantiated tile counts (theta tiles per half, even and odd l tiles) and
// whether the mixed form is built for them: at 3 theta tiles per half the
// mixed form's registers spill, and those jobs run the separate passes.
#define QD_SPLIT_CASES(X) \
  X(1, 1, 1, true) X(2, 1, 1, true) X(2, 2, 1, true) X(2, 2, 2, true) X(3, 3, 3, false)

bool leg_split_supported(int th, int te, int to, bool mixed) {
  switch (th * 64 + te * 8 + to) {
#define QD_SPLIT_OK(H, E, O, M) \
  case H * 64 + E * 8 + O: return !mixed || M;
    QD_SPLIT_CASES(QD_SPLIT_OK)
#undef QD_SPLIT_OK
    default: return false;
  }
}

size_t leg_split_lds_bytes(int th, int te, int to, int mixed) {
  const size_t f = (size_t)16 * ((size_t)th * lds_stride(te) + (size_t)th * lds_stride(to) +
                                 (size_t)te * lds_stride(th) + (size_t)to * lds_stride(th));
  // two-sweep form: 4 waves x th chunks x 4 k-steps x 64 lanes of 16-byte operands
  if (!mixed) return f * sizeof(double) + (th >= 3 ? (size_t)4 * th * 4 * 64 * 16 : 0);
  // compensation entries: 4 rows of 16-byte entries per 16 k, all columns
  const size_t x = (size_t)4 * 16 * ((size_t)th * te + (size_t)th * to + (size_t)te * th + (size_t)to * th);
  return x * 16 + f * sizeof(float);
}

template <int H, int E, int O, bool M>
static void launch_split_case(const StepArgs &a, int total_units, int mixed, size_t lds, int finish, hipStream_t s) {
  const dim3 grid(total_units), block(256);
  if constexpr (M) {
    if (mixed) {
      if (finish) hipLaunchKernelGGL((k_leg_split<H, E, O, true, true>), grid, block, lds, s, a);
      else hipLaunchKernelGGL((k_leg_split<H, E, O, false, true>), grid, block, lds, s, a);
      return;
    }
  }
  if (finish) hipLaunchKernelGGL((k_leg_split<H, E, O, true, false>), grid, block, lds, s, a);
  else hipLaunchKernelGGL((k_leg_split<H, E, O, false, false>), grid, block, lds, s, a);
}

void launch_leg_split(const StepArgs &a, int total_units, int th, int te, int to, int mixed, size_t lds, int finish,
                      hipStream_t s) {
  switch (th * 64 + te * 8 + to) {
#define QD_SPLIT_LAUNCH(H, E, O, M) \
  case H * 64 + E * 8 + O: launch_split_case<H, E, O, M>(a, total_units, mixed, lds, finish, s); break;
    QD_SPLIT_CASES(QD_SPLIT_LAUNCH)
#undef QD_SPLIT_LAUNCH
    default: break;   // excluded by leg_split_supported
  }
}

template <int H, int E, int O, bool M>
static void attrs_split_case(int lim) {
  const hipFuncAttribute at = hipFuncAttributeMaxDynamicSharedMemorySize;
  (void)hipFuncSetAttribute((const void *)k_leg_split<H, E, O, false, false>, at, lim);
  (void)hipFuncSetAttribute((const void *)k_leg_split<H, E, O, true, false>, at, lim);
  if constexpr (M) {
    (void)hipFuncSetAttribute((const void *)k_leg_split<H, E, O, false, true>, at, lim);
    (void)hipFuncSetAttribute((const void *)k_leg_split<H, E, O, true, true>, at, lim);
  }
}

void init_attrs_leg_split() {
#define QD_SPLIT_ATTR(H, E, O, M) attrs_split_case<H, E, O, M>(160 * 1024);
  QD_SPLIT_CASES(QD_SPLIT_ATTR)
#undef QD_SPLIT_ATTR
}

}  // namespace qd
