// flux_surface.h - the flux surface work of one step (kernels_flux.hip),
// also run as leading workgroups of the step's first FFT pass.
#pragma once
#include "qd_internal.h"
#include "qd_complex.h"

namespace qd {

// Surface values of one step into the flux window (the surface kernel's
// work): one thread per surface point s = (b nt + k) n1 + i of every flux job,
// u the work unit (workgroup) of the surface kind.
__device__ __forceinline__ void flux_unit(const StepArgs &a, int u) {
  const int q = locate_job(a.unit_base, a.n_jobs, u);
  const QD_G JobDev &jb = a.jobs[q];
  const int s = (u - a.unit_base[q]) * kBlock + threadIdx.x;
  // in the fused step sequence: only jobs whose last step is finishing now
  if (!jb.flux_on || s >= jb.flux_S || (a.gate && jb.counters[3] == 0)) return;
  const int n1 = jb.n1, nt = jb.nt, M = jb.M, ng = jb.flux_ng, half = ng / 2;
  const int i = s % n1, k = (s / n1) % nt, b = s / (n1 * nt);
  const QD_G double2 *p = jb.psi + ((long)b * nt + k) * M + (long)(jb.flux_ns - half) * n1 + i;
  const QD_G double2 *rph = jb.flux_rph + (long)k * n1 + i;
  const long rstride = (long)nt * n1;
  double2 d = make_double2(0.0, 0.0), v = make_double2(0.0, 0.0);
  for (int g = 0; g < ng; g++) {
    const double2 x = cmul(p[(long)g * n1], rph[g * rstride]);
    const double c = jb.flux_coef[g];
    d.x = fma(c, x.x, d.x);
    d.y = fma(c, x.y, d.y);
    if (g == half) v = x;
  }
  const int slot = jb.counters[0] % jb.flux_T;   // the step counter before this step's increment
  QD_G double2 *W = jb.flux_W + (long)slot * 2 * jb.flux_S;
  W[s] = v;
  W[jb.flux_S + s] = d;
}

// The surface work carried by the first FFT pass of the fused step sequence:
// its workgroups red_units .. red_units + flux_units - 1 (after the carried
// reduction's) write the previous step's surface values (gated); returns
// true for those, and the FFT work takes the rest.  The FFT pass reads and
// writes neither the surface rows nor the window, and the step counter and
// pending flag the surface work reads do not change during it.
__device__ __forceinline__ bool flux_carried(const StepArgs &a) {
  const int u = (int)blockIdx.x - a.red_units;
  if (u < 0 || u >= a.flux_units) return false;
  StepArgs f = a;
  f.unit_base = a.flux_base;
  f.gate = 1;
  flux_unit(f, u);
  return true;
}

}  // namespace qd
