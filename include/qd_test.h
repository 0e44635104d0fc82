/*
 * qd_test.h - diagnostic entry points of libqdamd, used by the test suite
 * to check single building blocks of the propagator on the device.
 * Each replaces one vendor routine the reference calls on its hot path.
 */
#ifndef QD_TEST_H
#define QD_TEST_H

#include <stdint.h>
#include "qd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Real GEMM C (rows x n) = A (rows x k) . B (k x n), all column-major, run
 * with the Legendre-transform MFMA kernel.  Replaces the cublasZgemm calls
 * with a zero-imaginary B (omegawavepacket.cu:212-218, :233-239). n <= 128. */
qd_status qd_probe_gemm(const double *A, int32_t rows, int32_t k, const double *B, int32_t n, double *C);

/* In-place unnormalised DFT of `count` contiguous complex pencils of length n
 * with the LDS Stockham kernel (forward exp(-2 pi i jk/n), inverse +).
 * Replaces cufftExecZ2Z (omegawavepacket.cu:242-269). */
qd_status qd_probe_fft(double *data, int32_t n, int32_t count, int32_t inverse);

/* Eigendecomposition of the n x n Coriolis matrix C_l over
 * omega = omega_min..omega_min+n-1 (host solver; replaces LAPACK dstev,
 * CoriolisMatrix.C:63-99).  evecs[alpha*n + i] = component i of vector alpha. */
qd_status qd_probe_coriolis(int32_t J, int32_t l, int32_t omega_min, int32_t n, double *evals, double *evecs);

/* Allocation-free variant of qd_probe_fft, with no counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  `data` is DEVICE memory holding
 * `count` contiguous complex pencils of length n (interleaved re, im), which
 * are transformed in place on `stream` (a hipStream_t; NULL = the default
 * stream); the call returns without waiting for them.  The twiddle table
 * lives in a static device array of 512 entries per GPU, so n must be at most
 * 512 (QD_ERR_CAPACITY otherwise) and factor into 2, 3, 5.  Every call on a
 * GPU rewrites that GPU's table: calls that overlap on one GPU race. */
qd_status qd_fft_pencils_noalloc(double *data, int32_t n, int32_t count, int32_t inverse, void *stream);

/* Plain device-to-device copy of `bytes` bytes, `repeats` times, timed with
 * HIP events; calibrates the HBM byte counters (FETCH_SIZE / WRITE_SIZE)
 * against a kernel of known traffic.  No counterpart in the reference. */
qd_status qd_probe_copy(int64_t bytes, int32_t repeats, double *us_per_copy);

/* HBM ceiling probes over `bytes` bytes, `repeats` passes timed with HIP
 * events: kind 1 copies (reads and writes `bytes`) with eight 16-byte loads
 * in flight per thread and nontemporal stores; kind 2 only reads.  Used to
 * quote kernels against a measured ceiling and to calibrate the byte
 * counters.  No counterpart in the reference. */
qd_status qd_probe_bandwidth(int64_t bytes, int32_t repeats, int32_t kind, double *us_per_pass);

/* Which form of the Legendre transforms a job built from `in` would run (host
 * only: validates the inputs, touches no device): 0 the separate forward and
 * backward passes, 1 the fused step (backward transform of one step with the
 * forward transform of the next), 2 the fused step with parity-split products
 * (the theta grid and tables symmetric under x -> -x).  No counterpart in the
 * reference. */
qd_status qd_probe_legendre_form(const qd_inputs *in, int32_t *form);

/* How many read-write passes over the FBR array the FFT work of a job built
 * from `in` would take (host only, like qd_probe_legendre_form): 3 (R pass
 * with the first rotational . Coriolis . rotational half step, r pass with
 * the kinetic phase, R pass with the second half step) or 2 (r pass with the
 * first half step and the r kinetic factor, R pass with the R kinetic factor
 * and the second half step; compile-time lengths, up to 5 omega blocks;
 * QD_FFT_TWO_PASS=0 in the environment turns it off).  No counterpart in the
 * reference. */
qd_status qd_probe_fft_passes(const qd_inputs *in, int32_t *passes);

#ifdef __cplusplus
}
#endif

#endif /* QD_TEST_H */
