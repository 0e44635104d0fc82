// qd_complex.h - double2 complex arithmetic used by the device kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace qd {

__host__ __device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * b + c
__host__ __device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
  return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
__host__ __device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ __forceinline__ double cnorm(double2 a) { return a.x * a.x + a.y * a.y; }
// Re(conj(a) * b)
__host__ __device__ __forceinline__ double cdot_re(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }
// multiply by -i (forward) or +i (inverse)
__host__ __device__ __forceinline__ double2 cmul_mi(double2 a) { return make_double2(a.y, -a.x); }
__host__ __device__ __forceinline__ double2 cmul_pi(double2 a) { return make_double2(-a.y, a.x); }

}  // namespace qd
