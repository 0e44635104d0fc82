// step_reduce.h - one workgroup of the per-step observable reduction, shared
// by the reduction kernel and the first FFT pass of the fused step sequence
// (the forward R pass, or the r pass with the first half step in the two-pass
// form), which carries the previous step's reduction as leading workgroups.
#pragma once

#include "qd_internal.h"

namespace qd {

// Workgroup `unit` of the reduction (256 threads; red holds >= 4 doubles).
// Workgroup (b, f) of a job sums all partials of field f of omega block b in
// a fixed order (thread t takes elements t, t + 256, ... in turn; then a
// fixed wave and workgroup sum) and writes that field of the step record: no
// workgroup waits for another, so no device-scope fence is needed (each of
// which writes back an XCD's L2 on this chip).  Workgroup (0, 0) sets the
// pending-advance flag; one workgroup of the next Legendre kernel advances
// the step counters (advance_counters), so no workgroup here can read a
// counter that another has already advanced.
// Gated (the fused step sequence): only jobs whose backward transform was
// pending, i.e. ran in the fused kernel just before, have a step to record.
// Ungated (after a separate backward pass): every job, and workgroup (0, 0)
// clears the pending flag (no workgroup of an ungated launch reads it).
// Jobs stepped with the fused kernel take their potential-energy partials
// from its workgroups (fused or backward-only form), others from those of
// the backward Legendre pass.
__device__ __forceinline__ void reduce_unit(const StepArgs &a, int unit, double *red) {
  const int job = locate_job(a.unit_base, a.n_jobs, unit);
  const JobDev &jb = a.jobs[job];
  const int u = unit - a.unit_base[job];
  if (a.gate && jb.counters[3] == 0) return;   // uniform over the workgroup
  const int nO = jb.nO;
  const int b = u / kRedFields, f = u % kRedFields;
  const QD_G double *src;
  long n;
  if (f == 0 && jb.fft2) {   // kinetic energy (two-pass form): [b][units] of the inverse R pass
    n = jb.units[KK_FFT_RI];
    src = jb.part_ke + (long)b * n;
  } else if (f == 0) {   // kinetic energy: FFT_r units ordered by block, slab, column group
    const int gcols = jb.n1 / jb.fft_cols;
    long off = 0;
    for (int q = 0; q < b; q++) off += (long)(jb.L - (jb.omin + q)) * gcols;
    src = jb.part_ke + off;
    n = (long)(jb.L - (jb.omin + b)) * gcols;
  } else if (f < 6) {
    // second FBR half step partials of the inverse-R pass: [b][5][units] =
    // Coriolis energy, rotational energy, norm, autocorrelation re / im
    n = jb.units[KK_FFT_RI];
    src = jb.part_cor + (5L * b + (f - 1)) * n;
  } else {        // potential energy partials of the backward Legendre epilogue: [b][group]
    n = jb.leg_step ? jb.step_groups : jb.gemm_bwd_groups;   // fused kernel (both forms) or backward pass
    src = jb.part_obs + (long)b * n;
  }
  // eight loads in flight per thread, summed in element order
  double acc = 0.0;
  for (long q0 = threadIdx.x; q0 < n; q0 += 8L * kBlock) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const long q = q0 + (long)k * kBlock;
      v[k] = q < n ? src[q] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) acc += v[k];
  }
  acc = wave_sum64(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double sum = (red[0] + red[1]) + (red[2] + red[3]);
    const int slot = jb.counters[1] % jb.ring;
    const double area = jb.dr1 * jb.dr2;
    QD_G double *rec = jb.obs_ring + ((long)slot * nO + b) * 5;
    QD_G double *acf = (QD_G double *)(jb.acf_ring + (long)slot * nO + b);
    switch (f) {
      case 0: rec[2] = sum * area / (double)jb.M; break;   // kinetic
      case 1: rec[4] = sum * area; break;                  // Coriolis
      case 2: rec[3] = sum * area; break;                  // rotational
      case 3: rec[0] = sum * area; break;                  // norm
      case 4: acf[0] = sum * area; break;
      case 5: acf[1] = sum * area; break;
      default: rec[1] = sum * area; break;                 // potential
    }
    if (u == 0) {
      jb.counters[2] = 1;
      if (!a.gate) jb.counters[3] = 0;
    }
  }
}

// The reduction carried by the first FFT pass of the fused step sequence: its
// first a.red_units workgroups reduce (gated) the previous step; returns true
// for those, and the FFT work takes the rest (unit = blockIdx.x - red_units,
// less the carried flux work's units).
__device__ __forceinline__ bool reduce_carried(const StepArgs &a, double *red) {
  if ((int)blockIdx.x >= a.red_units) return false;
  StepArgs r = a;
  r.unit_base = a.red_base;
  r.gate = 1;
  reduce_unit(r, blockIdx.x, red);
  return true;
}

}  // namespace qd
